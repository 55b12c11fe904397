// fixed_real.hip -- compile-time plans, real FP32 (half-length cores): 2048 (1024 = 16 16 4).
// FixedReal<type, N, threads per transform, transforms per block,
//           Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//           variant (0 = default), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedReal<float, 1024, 64, 4, Sched<0, 0, 0>, 0, 16, 16, 4> r2048a;
const FixedReal<float, 1024, 64, 4, Sched<-1, 0, 0>, 1, 16, 16, 4> r2048b;
const FixedReal<float, 1024, 64, 4, Sched<0, 4, 0>, 2, 16, 16, 4> r2048c;
const FixedReal<float, 1024, 64, 8, Sched<0, 0, 0>, 3, 16, 16, 4> r2048d;
const FixedReal<float, 1024, 64, 2, Sched<0, 0, 0>, 4, 16, 16, 4> r2048e;
const FixedReal<float, 1024, 128, 4, Sched<0, 0, 0>, 5, 16, 16, 4> r2048f;
}  // namespace
}  // namespace sfft
