// fixed_single0.hip -- compile-time plans, complex FP32, single pass: part
// 1 of 4 of the fast lengths 2^a 3^b 5^c in [64, 19200] that have no
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
const FixedC2C<float, 64, 4, 64, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 4> c64;  // 4834 -> 5131
const FixedC2C<float, 250, 10, 25, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 10> c250;  // 4122 -> 4544
const FixedC2C<float, 324, 36, 7, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6> c324;  // 4004 -> 4571
const FixedC2C<float, 500, 50, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 5> c500;  // 4545 -> 4950
const FixedC2C<float, 729, 81, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9> c729;  // 4486 -> 4818
const FixedC2C<float, 972, 108, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9> c972;  // 4137 -> 4861
const FixedC2C<float, 1296, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 9> c1296;  // 3627 -> 5000
const FixedC2C<float, 1536, 128, 2, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 16, 16, 6> c1536;  // 4514 -> 4749
const FixedC2C<float, 1944, 216, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6, 6> c1944;  // 3116 -> 4305
const FixedC2C<float, 2250, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 10> c2250;  // 3852 -> 4323
const FixedC2C<float, 2592, 96, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 9, 9> c2592;  // 3649 -> 4711
const FixedC2C<float, 3072, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 16, 16, 12> c3072;  // 4616 -> 4856
const FixedC2C<float, 3645, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 5> c3645;  // 2636 -> 4313
const FixedC2C<float, 4320, 160, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 15, 9> c4320;  // 2651 -> 4137
const FixedC2C<float, 4800, 160, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 15, 10> c4800;  // 2699 -> 4347
const FixedC2C<float, 5400, 360, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 8, 3> c5400;  // 2795 -> 3300
const FixedC2C<float, 6000, 400, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 16, 15> c6000;  // 3631 -> 4166
const FixedC2C<float, 6400, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 8> c6400;  // 3649 -> 4520
const FixedC2C<float, 7200, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 15, 15> c7200;  // 3394 -> 4331
const FixedC2C<float, 8000, 320, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 10> c8000;  // 2433 -> 3788
const FixedC2C<float, 9216, 768, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 6, 6> c9216;  // 2393 -> 3904
const FixedC2C<float, 10240, 683, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 8, 5> c10240;  // 2426 -> 3817
const FixedC2C<float, 11664, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 9, 9, 9> c11664;  // 1962 -> 3545
const FixedC2C<float, 12960, 480, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 9, 9, 5> c12960;  // 1917 -> 3446
const FixedC2C<float, 19200, 640, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 10, 10, 6> c19200;  // 1818 -> 2153
}  // namespace
}  // namespace sfft
