// fixed_single1.hip -- compile-time plans, complex FP32, single pass: part
// 2 of 4 of the fast lengths 2^a 3^b 5^c in [64, 19200] that have no
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
const FixedC2C<float, 125, 9, 28, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 5> c125;  // 4060 -> 4314
const FixedC2C<float, 270, 23, 11, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 3> c270;  // 4145 -> 4464
const FixedC2C<float, 384, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 12> c384;  // 4657 -> 4974
const FixedC2C<float, 540, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6> c540;  // 4116 -> 4361
const FixedC2C<float, 800, 32, 8, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 25> c800;  // 4328 -> 4811
const FixedC2C<float, 1000, 100, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10> c1000;  // 4081 -> 4604
const FixedC2C<float, 1440, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 10> c1440;  // 3897 -> 4952
const FixedC2C<float, 1600, 54, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 10, 5> c1600;  // 4194 -> 4438
const FixedC2C<float, 2025, 81, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 9, 9> c2025;  // 3247 -> 4518
const FixedC2C<float, 2400, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 10> c2400;  // 4077 -> 4806
const FixedC2C<float, 2700, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 12> c2700;  // 4227 -> 4513
const FixedC2C<float, 3125, 125, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 25, 5> c3125;  // 4028 -> 4646
const FixedC2C<float, 3888, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6, 6> c3888;  // 2701 -> 4333
const FixedC2C<float, 4374, 486, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 6> c4374;  // 2952 -> 4376
const FixedC2C<float, 4860, 540, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 6> c4860;  // 2075 -> 3879
const FixedC2C<float, 5625, 375, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 15, 15> c5625;  // 3456 -> 3886
const FixedC2C<float, 6075, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 3, 3, 3> c6075;  // 2612 -> 3175
const FixedC2C<float, 6480, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 9, 9, 5> c6480;  // 2106 -> 4432
const FixedC2C<float, 7290, 810, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 9> c7290;  // 2494 -> 4064
const FixedC2C<float, 8100, 900, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 9, 9> c8100;  // 2428 -> 3909
const FixedC2C<float, 9600, 640, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 8, 5> c9600;  // 2347 -> 3779
const FixedC2C<float, 10368, 864, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12, 6> c10368;  // 2554 -> 3779
const FixedC2C<float, 12000, 480, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 15> c12000;  // 2418 -> 3702
const FixedC2C<float, 14400, 960, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15, 4> c14400;  // 2646 -> 3489
}  // namespace
}  // namespace sfft
