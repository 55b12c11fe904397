// fixed_real.hip -- compile-time plans, real FP32 (half-length cores): 2048 (1024 = 16 16 4).
// FixedReal<type, N, threads per transform, transforms per block,
//           Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//           variant (0 = default), radices...>
#include "fixed.h"

namespace sfft {
namespace {
// default: two frames per block (4.49 / 4.15 TB/s forward / inverse on
// inputs the last-level cache does not hold; four per block 4.41 / 3.89,
// eight 3.51 / 3.71, 128 threads per frame 3.61 / 3.54; two stages of radix
// 32 on 32 threads per frame, 2 to 8 frames per block: forward 2.8-3.5 TB/s)
const FixedReal<float, 1024, 64, 2, Sched<0, 0, 0>, 0, 16, 16, 4> r2048e;
const FixedReal<float, 1024, 64, 4, Sched<-1, 0, 0>, 1, 16, 16, 4> r2048b;
const FixedReal<float, 1024, 64, 4, Sched<0, 4, 0>, 2, 16, 16, 4> r2048c;
const FixedReal<float, 1024, 64, 4, Sched<0, 0, 0>, 3, 16, 16, 4> r2048a;
const FixedReal<float, 1024, 64, 2, Sched<0, 0, 0, 0, 5>, 4, 16, 16, 4> r2048p5;
const FixedReal<float, 1024, 64, 2, Sched<0, 0, 0, 0, 3>, 5, 16, 16, 4> r2048p3;
}  // namespace
}  // namespace sfft
