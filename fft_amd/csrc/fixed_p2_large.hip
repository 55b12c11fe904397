// fixed_p2_large.hip -- compile-time plans, complex FP32: 16384, 8192.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
// default: three stages of radix 32/32/16 on 512 threads (256 registers per
// thread, two exchanges instead of three): 4.35 TB/s on inputs the last-level
// cache does not hold, against 4.08-4.21 for four stages of 16/16/16/4 on
// 1024 threads
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 1>, 0, 32, 32, 16> c16384c;
const FixedC2C<float, 16384, 512, 1, Sched<-1, 0, 0>, 3, 32, 32, 16> c16384d;
const FixedC2C<float, 16384, 1024, 1, Sched<0, 0, 0, 1>, 2, 16, 16, 16, 4> c16384a;
const FixedC2C<float, 16384, 1024, 1, Sched<-1, 0, 0>, 1, 16, 16, 16, 4> c16384b;
const FixedC2C<float, 8192, 512, 1, Sched<-1, 4, 0>, 0, 16, 16, 8, 4> c8192a;
const FixedC2C<float, 8192, 512, 1, Sched<0, 4, 0>, 1, 16, 16, 8, 4> c8192b;
}  // namespace
}  // namespace sfft
