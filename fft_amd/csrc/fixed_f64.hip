// fixed_f64.hip -- compile-time plans, FP64: complex 256...8192; real 2048, 4096, 8192
// (16-byte values: 225-254 VGPRs at 16 elements per thread, two waves per
// SIMD).  Measured against the runtime-planned pass on 2 GiB batches:
// 1.45-2.1x (5.0-5.7 vs 2.3-3.6 TB/s), and 14 vs 25 us for one 4096 round trip.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads>,
//          variant (0 = default, the fastest measured), radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<double, 4096, 256, 1, Sched<-1, 0, 0>, 0, 16, 16, 16> d4096a;
const FixedC2C<double, 4096, 256, 1, Sched<0, 0, 0>, 1, 16, 16, 16> d4096b;
const FixedC2C<double, 4096, 512, 1, Sched<-1, 0, 0>, 2, 16, 16, 16> d4096c;
const FixedC2C<double, 1024, 64, 2, Sched<-1, 0, 0>, 1, 16, 16, 4> d1024a;
const FixedC2C<double, 1024, 64, 4, Sched<-1, 0, 0>, 0, 16, 16, 4> d1024b;
const FixedC2C<double, 2048, 128, 1, Sched<-1, 0, 0>, 1, 16, 16, 8> d2048a;
const FixedC2C<double, 2048, 128, 2, Sched<-1, 0, 0>, 0, 16, 16, 8> d2048b;
const FixedC2C<double, 8192, 512, 1, Sched<-1, 0, 0>, 0, 16, 16, 8, 4> d8192a;
const FixedC2C<double, 8192, 512, 1, Sched<0, 0, 0>, 1, 16, 16, 8, 4> d8192b;
const FixedC2C<double, 256, 16, 8, Sched<-1, 0, 0>, 0, 16, 16> d256a;
const FixedC2C<double, 512, 32, 4, Sched<-1, 0, 0>, 0, 16, 8, 4> d512a;
// real FP64 (half-length cores): forward 1.36-1.84x, inverse 1.23-1.38x the
// runtime-planned pass on 2 GiB batches
const FixedReal<double, 1024, 64, 2, Sched<-1, 0, 0>, 0, 16, 16, 4> dr2048a;
const FixedReal<double, 2048, 128, 1, Sched<-1, 0, 0>, 0, 16, 16, 8> dr4096a;
const FixedReal<double, 4096, 256, 1, Sched<-1, 0, 0>, 0, 16, 16, 16> dr8192a;
}  // namespace
}  // namespace sfft
