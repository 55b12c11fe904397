// fixed_p2_small.hip -- compile-time plans, complex FP32: 1024, 512, 256.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 1024, 64, 4, Sched<0, 4, 0>, 0, 16, 16, 4> c1024a;
const FixedC2C<float, 1024, 64, 4, Sched<-1, 0, 0>, 1, 16, 16, 4> c1024b;
const FixedC2C<float, 512, 32, 8, Sched<0, 4, 0>, 0, 16, 8, 4> c512a;
const FixedC2C<float, 512, 32, 8, Sched<-1, 0, 0>, 1, 16, 8, 4> c512b;
const FixedC2C<float, 256, 16, 16, Sched<-1, 0, 0>, 0, 16, 16> c256a;
const FixedC2C<float, 256, 16, 16, Sched<0, 4, 0>, 1, 16, 16> c256b;
}  // namespace
}  // namespace sfft
