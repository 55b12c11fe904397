// fixed_single2.hip -- compile-time plans, complex FP32, single pass: part
// 3 of 4 of the fast lengths 2^a 3^b 5^c in [64, 19200] that have no
// hand-tuned kernel.  For each length two candidates were measured against the
// runtime-planned pass (batched launches of about 256 MB, cold inputs): the
// radices the runtime cost model picks at 16 elements per thread and at 32
// (radix 32 allowed), threads per transform as that model computes them, and
// about 256 threads per block.  A length is listed only where its faster
// candidate beat the runtime pass by 5 % or more (the rate pair in GB/s, runtime
// -> compiled, follows each line).  Lengths where the candidates lost (up to
// 2.4x slower, mostly lengths above 10000 at 600 to 1000 threads per
// transform) keep the runtime pass.  Lines whose schedule ends in 3 use
// non-temporal input loads and output stores (Sched NT): 800, 1536, 1728,
// 2160, 2430, 2560, 3072, 7200, 7680, 12288, 12800, 17280 and 19200 measured
// 7-18 % faster with them, others up to 2.5x slower.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first,
//          LDS pad shift>, variant, radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 162, 18, 14, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 2> c162;  // 4076 -> 4777
const FixedC2C<float, 288, 24, 10, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 2> c288;  // 4478 -> 4880
const FixedC2C<float, 405, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 5> c405;  // 4007 -> 4673
const FixedC2C<float, 576, 48, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6, 6> c576;  // 4426 -> 4888
const FixedC2C<float, 810, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9> c810;  // 4286 -> 4849
const FixedC2C<float, 1080, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 6> c1080;  // 4323 -> 4626
const FixedC2C<float, 1458, 162, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 2> c1458;  // 3096 -> 4339
const FixedC2C<float, 1620, 180, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 9> c1620;  // 3928 -> 4622
const FixedC2C<float, 2160, 180, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 15, 12, 12> c2160;  // 4366 -> 4645
const FixedC2C<float, 2430, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 10, 9, 9, 3> c2430;  // 2060 -> 3701
const FixedC2C<float, 2880, 96, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 15, 6> c2880;  // 4489 -> 4919
const FixedC2C<float, 3240, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6, 6> c3240;  // 2167 -> 3050
const FixedC2C<float, 4000, 160, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 5> c4000;  // 2794 -> 4690
const FixedC2C<float, 4500, 188, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 15, 12> c4500;  // 2461 -> 3238
const FixedC2C<float, 5120, 171, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 32, 5> c5120;  // 2876 -> 4420
const FixedC2C<float, 5760, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 15, 12> c5760;  // 3084 -> 4608
const FixedC2C<float, 6144, 205, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 32, 6> c6144;  // 3192 -> 4298
const FixedC2C<float, 6561, 729, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 9> c6561;  // 2575 -> 4060
const FixedC2C<float, 7680, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 16, 15> c7680;  // 3422 -> 4318
const FixedC2C<float, 8640, 720, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12, 5> c8640;  // 2257 -> 3813
const FixedC2C<float, 10000, 1000, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10, 10> c10000;  // 2554 -> 3956
const FixedC2C<float, 10800, 720, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15, 3> c10800;  // 2486 -> 3622
const FixedC2C<float, 12288, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 32, 12> c12288;  // 2616 -> 3760
const FixedC2C<float, 15360, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 32, 15> c15360;  // 2775 -> 3788
}  // namespace
}  // namespace sfft
