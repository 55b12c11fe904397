// fixed_p2_large.hip -- compile-time plans, complex FP32: 16384, 8192.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
// default: three stages of radix 32/32/16 on 512 threads (256 registers per
// thread, two exchanges instead of three) with one LDS pad element per 32:
// the first stage writes 32-element runs, and the pad starts each run 8 bytes
// further along the banks.  On inputs the last-level cache does not hold:
// 4.59 TB/s, against 4.35 with one pad per 16 and 4.03-4.31 for four stages
// of 16/16/16/4 on 1024 threads.
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 0, 5>, 0, 32, 32, 16> c16384f;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 1, 5>, 1, 32, 32, 16> c16384e;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 1, 4>, 2, 32, 32, 16> c16384c;
const FixedC2C<float, 16384, 1024, 1, Sched<0, 0, 0, 1, 4>, 3, 16, 16, 16, 4> c16384a;
const FixedC2C<float, 16384, 1024, 1, Sched<-1, 0, 0>, 4, 16, 16, 16, 4> c16384b;
const FixedC2C<float, 16384, 512, 1, Sched<-1, 0, 0, 0, 5>, 5, 32, 32, 16> c16384h;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 0, 6>, 6, 32, 32, 16> c16384i;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 1, 0, 5>, 7, 32, 32, 16> c16384j;
// non-temporal input loads / output stores / both measured slower on the
// bench's 1 GiB batches (4.19 / 4.27 / 4.33 against 4.59-4.63 TB/s)
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 0, 5, 1>, 8, 32, 32, 16> c16384nl;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 0, 5, 2>, 9, 32, 32, 16> c16384ns;
const FixedC2C<float, 16384, 512, 1, Sched<0, 0, 0, 0, 5, 3>, 10, 32, 32, 16> c16384n;
const FixedC2C<float, 8192, 512, 1, Sched<-1, 4, 0>, 0, 16, 16, 8, 4> c8192a;
const FixedC2C<float, 8192, 512, 1, Sched<0, 4, 0>, 1, 16, 16, 8, 4> c8192b;
}  // namespace
}  // namespace sfft
