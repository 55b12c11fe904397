// fixed_f64_single2.hip -- compile-time plans, complex FP64, single pass:
// part 3 of 4 of the fast lengths 2^a 3^b 5^c in [64, 10000] that have no
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
const FixedC2C<double, 80, 6, 59, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 5> d80;  // 4602 -> 5211
const FixedC2C<double, 100, 10, 47, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10> d100;  // 4811 -> 5431
const FixedC2C<double, 144, 12, 33, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12> d144;  // 4819 -> 5277
const FixedC2C<double, 180, 15, 26, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12> d180;  // 4796 -> 5142
const FixedC2C<double, 225, 15, 21, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15> d225;  // 4757 -> 5130
const FixedC2C<double, 288, 24, 10, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 2> d288;  // 4530 -> 5258
const FixedC2C<double, 360, 24, 10, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 8, 3> d360;  // 4331 -> 5051
const FixedC2C<double, 405, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 5> d405;  // 3265 -> 4926
const FixedC2C<double, 486, 54, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6> d486;  // 3369 -> 4980
const FixedC2C<double, 600, 40, 6, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 8, 5> d600;  // 4173 -> 5008
const FixedC2C<double, 675, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 3> d675;  // 4215 -> 4812
const FixedC2C<double, 768, 52, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 3> d768;  // 4528 -> 5043
const FixedC2C<double, 900, 60, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 4> d900;  // 4344 -> 4751
const FixedC2C<double, 1080, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 6> d1080;  // 3869 -> 5045
const FixedC2C<double, 1215, 135, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 9, 9> d1215;  // 2488 -> 4020
const FixedC2C<double, 1350, 113, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 6> d1350;  // 3866 -> 4322
const FixedC2C<double, 1536, 128, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 6> d1536;  // 4521 -> 5154
const FixedC2C<double, 1800, 120, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 8> d1800;  // 4300 -> 4790
const FixedC2C<double, 2000, 134, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 5, 5, 5> d2000;  // 2974 -> 4425
const FixedC2C<double, 2250, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 10> d2250;  // 3522 -> 4699
const FixedC2C<double, 2500, 250, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 5, 5> d2500;  // 3343 -> 4860
const FixedC2C<double, 2880, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 12> d2880;  // 4298 -> 5101
const FixedC2C<double, 3125, 209, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 5, 5, 5, 5, 5> d3125;  // 3022 -> 3273
const FixedC2C<double, 3456, 288, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6, 6, 6> d3456;  // 2546 -> 4870
const FixedC2C<double, 3840, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 15> d3840;  // 4422 -> 4926
const FixedC2C<double, 4374, 486, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 6> d4374;  // 2296 -> 3696
const FixedC2C<double, 5000, 500, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10, 5> d5000;  // 2641 -> 4245
const FixedC2C<double, 5832, 486, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 6, 6, 6, 3, 3, 3> d5832;  // 2199 -> 3515
const FixedC2C<double, 6480, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 3, 3, 3> d6480;  // 2410 -> 3318
}  // namespace
}  // namespace sfft
