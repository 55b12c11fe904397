// fixed_p2_large.hip -- compile-time plans, complex FP32: 16384, 8192.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 16384, 1024, 1, Sched<0, 0, 0, 1>, 0, 16, 16, 16, 4> c16384a;
const FixedC2C<float, 16384, 1024, 1, Sched<-1, 0, 0>, 1, 16, 16, 16, 4> c16384b;
// three stages of radix 32/32/16 on 512 threads (256 registers per thread)
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 1>, 2, 32, 32, 16> c16384c;
const FixedC2C<float, 16384, 512, 1, Sched<-1, 0, 0>, 3, 32, 32, 16> c16384d;
const FixedC2C<float, 8192, 512, 1, Sched<-1, 4, 0>, 0, 16, 16, 8, 4> c8192a;
const FixedC2C<float, 8192, 512, 1, Sched<0, 4, 0>, 1, 16, 16, 8, 4> c8192b;
}  // namespace
}  // namespace sfft
