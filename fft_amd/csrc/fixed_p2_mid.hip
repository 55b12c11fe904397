// fixed_p2_mid.hip -- compile-time plans, complex FP32: 4096, 2048.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 4096, 256, 1, Sched<-1, 4, 0>, 0, 16, 16, 16> c4096a;
const FixedC2C<float, 4096, 256, 1, Sched<-2, 0, 0>, 1, 16, 16, 16> c4096b;
const FixedC2C<float, 2048, 128, 2, Sched<-1, 0, 0>, 0, 16, 16, 8> c2048a;
const FixedC2C<float, 2048, 128, 2, Sched<0, 4, 0>, 1, 16, 16, 8> c2048b;
}  // namespace
}  // namespace sfft
