// fixed_single3.hip -- compile-time plans, complex FP32, single pass: part
// 4 of 4 of the fast lengths 2^a 3^b 5^c in [64, 19200] that have no
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
const FixedC2C<float, 243, 27, 9, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 3> c243;  // 4215 -> 4826
const FixedC2C<float, 300, 13, 19, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 12> c300;  // 4325 -> 4616
const FixedC2C<float, 486, 54, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6> c486;  // 3892 -> 4497
const FixedC2C<float, 648, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6> c648;  // 3740 -> 4407
const FixedC2C<float, 864, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6> c864;  // 4476 -> 4764
const FixedC2C<float, 1215, 135, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 9, 9> c1215;  // 3274 -> 4291
const FixedC2C<float, 1500, 150, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10, 10> c1500;  // 3823 -> 4446
const FixedC2C<float, 1728, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 12, 12, 12> c1728;  // 4096 -> 4548
const FixedC2C<float, 2187, 243, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 3> c2187;  // 3574 -> 4633
const FixedC2C<float, 2560, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 16, 16, 10> c2560;  // 4306 -> 4838
const FixedC2C<float, 2916, 324, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6, 6> c2916;  // 2260 -> 3896
const FixedC2C<float, 3456, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 12, 9> c3456;  // 2393 -> 4102
const FixedC2C<float, 4050, 338, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 6, 3> c4050;  // 2410 -> 2933
const FixedC2C<float, 4608, 192, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 12, 12> c4608;  // 2702 -> 4612
const FixedC2C<float, 5184, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6, 6> c5184;  // 3114 -> 4454
const FixedC2C<float, 5832, 648, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9, 6> c5832;  // 2296 -> 3741
const FixedC2C<float, 6250, 250, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 25, 10> c6250;  // 2927 -> 3734
const FixedC2C<float, 6912, 576, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 6, 6> c6912;  // 2369 -> 4112
const FixedC2C<float, 7776, 864, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 9, 6> c7776;  // 2443 -> 4071
const FixedC2C<float, 8748, 972, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9, 9> c8748;  // 2518 -> 3752
const FixedC2C<float, 10125, 675, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 15, 3> c10125;  // 2434 -> 2876
const FixedC2C<float, 11520, 768, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 15, 3> c11520;  // 2508 -> 3740
const FixedC2C<float, 12800, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 25, 16> c12800;  // 2398 -> 3399
const FixedC2C<float, 17280, 576, 1, Sched<0, 0, 0, 0, kLdsPadShift, 3>, 0, 32, 15, 6, 6> c17280;  // 1886 -> 2038
}  // namespace
}  // namespace sfft
