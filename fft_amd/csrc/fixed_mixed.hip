// fixed_mixed.hip -- compile-time plans, complex FP32: 1920 = 16 12 10, 3000 = 12 10 25.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 1920, 160, 2, Sched<0, 0, 0>, 0, 16, 12, 10> c1920a;
const FixedC2C<float, 1920, 160, 2, Sched<0, 4, 0>, 1, 16, 12, 10> c1920b;
const FixedC2C<float, 1920, 160, 2, Sched<-1, 0, 0>, 2, 16, 12, 10> c1920c;
const FixedC2C<float, 3000, 250, 1, Sched<0, 0, 0>, 0, 12, 10, 25> c3000a;
const FixedC2C<float, 3000, 250, 1, Sched<-1, 0, 0>, 1, 12, 10, 25> c3000b;
}  // namespace
}  // namespace sfft
