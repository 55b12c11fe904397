// fixed_f64_single1.hip -- compile-time plans, complex FP64, single pass:
// part 2 of 4 of the fast lengths 2^a 3^b 5^c in [64, 10000] that have no
// hand-tuned kernel.  Each length was measured against the runtime-planned
// pass (batched launches of about 256 MB, cold inputs) with the runtime
// planner's radices (16 elements per thread, radix 16 at most: bitwise equal
// to it), at its threads per block and at twice as many transforms per
// block.  A length is listed only where the faster candidate beat the runtime
// pass by 5 % or more (the rate pair in GB/s, runtime -> compiled, follows
// each line); the others keep the runtime pass.
// FixedC2C<type, N, threads per transform, transforms per block,
//          Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first,
//          LDS pad shift>, variant, radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<double, 75, 5, 63, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 5> d75;  // 4390 -> 4986
const FixedC2C<double, 96, 8, 49, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6> d96;  // 4943 -> 5302
const FixedC2C<double, 135, 15, 17, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 9> d135;  // 4139 -> 5204
const FixedC2C<double, 162, 18, 14, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 2> d162;  // 3652 -> 5037
const FixedC2C<double, 216, 18, 22, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 6, 6, 6> d216;  // 4249 -> 5098
const FixedC2C<double, 270, 23, 11, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 3> d270;  // 4032 -> 5023
const FixedC2C<double, 324, 36, 7, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6> d324;  // 3896 -> 5266
const FixedC2C<double, 400, 27, 9, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 5, 5> d400;  // 4352 -> 4954
const FixedC2C<double, 480, 32, 8, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 2> d480;  // 4552 -> 5058
const FixedC2C<double, 576, 48, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6, 6> d576;  // 4349 -> 5126
const FixedC2C<double, 648, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6> d648;  // 3423 -> 4962
const FixedC2C<double, 750, 75, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10, 5> d750;  // 3544 -> 4991
const FixedC2C<double, 864, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6> d864;  // 4310 -> 5162
const FixedC2C<double, 1000, 100, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10> d1000;  // 3752 -> 5093
const FixedC2C<double, 1200, 80, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 5> d1200;  // 4309 -> 4921
const FixedC2C<double, 1296, 144, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 9> d1296;  // 2830 -> 4665
const FixedC2C<double, 1500, 150, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10, 10> d1500;  // 3052 -> 4808
const FixedC2C<double, 1728, 144, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12> d1728;  // 3442 -> 5003
const FixedC2C<double, 1944, 216, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6, 6> d1944;  // 3074 -> 4681
const FixedC2C<double, 2187, 243, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 3> d2187;  // 2908 -> 4403
const FixedC2C<double, 2430, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 3> d2430;  // 1859 -> 3908
const FixedC2C<double, 2700, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 12> d2700;  // 3818 -> 4843
const FixedC2C<double, 3072, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 12> d3072;  // 4416 -> 5120
const FixedC2C<double, 3375, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 15> d3375;  // 4020 -> 4752
const FixedC2C<double, 3750, 250, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 5, 5, 5, 2> d3750;  // 3196 -> 3582
const FixedC2C<double, 4320, 360, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6, 5> d4320;  // 2494 -> 4291
const FixedC2C<double, 4860, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6, 3, 3> d4860;  // 2206 -> 2559
const FixedC2C<double, 5760, 384, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 8, 3> d5760;  // 2612 -> 4162
const FixedC2C<double, 6400, 427, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 5, 5> d6400;  // 2657 -> 3747
const FixedC2C<double, 7680, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 15, 2> d7680;  // 2946 -> 3905
}  // namespace
}  // namespace sfft
