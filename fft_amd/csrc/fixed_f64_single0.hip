// fixed_f64_single0.hip -- compile-time plans, complex FP64, single pass:
// part 1 of 4 of the fast lengths 2^a 3^b 5^c in [64, 10000] that have no
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
const FixedC2C<double, 72, 6, 66, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 6> d72;  // 4890 -> 5164
const FixedC2C<double, 90, 10, 50, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9> d90;  // 4440 -> 5230
const FixedC2C<double, 125, 9, 28, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 5, 5, 5> d125;  // 3999 -> 4671
const FixedC2C<double, 160, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 10> d160;  // 4774 -> 5280
const FixedC2C<double, 200, 14, 18, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 8, 5, 5> d200;  // 3815 -> 4750
const FixedC2C<double, 250, 25, 19, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 5, 5> d250;  // 3991 -> 5107
const FixedC2C<double, 320, 22, 14, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 5, 4> d320;  // 4415 -> 5112
const FixedC2C<double, 384, 26, 9, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 8, 3> d384;  // 4564 -> 5108
const FixedC2C<double, 450, 30, 8, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 2> d450;  // 4391 -> 4843
const FixedC2C<double, 540, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6> d540;  // 3980 -> 5028
const FixedC2C<double, 640, 43, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 8, 5> d640;  // 4314 -> 5017
const FixedC2C<double, 729, 81, 6, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9> d729;  // 3407 -> 4723
const FixedC2C<double, 810, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9> d810;  // 3401 -> 4649
const FixedC2C<double, 972, 108, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9> d972;  // 3232 -> 4779
const FixedC2C<double, 1152, 96, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 6> d1152;  // 4161 -> 5215
const FixedC2C<double, 1280, 86, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 5> d1280;  // 4241 -> 5105
const FixedC2C<double, 1458, 162, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 2> d1458;  // 2525 -> 4070
const FixedC2C<double, 1620, 180, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 9> d1620;  // 3118 -> 4507
const FixedC2C<double, 1920, 128, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 8> d1920;  // 4403 -> 5012
const FixedC2C<double, 2160, 180, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 12> d2160;  // 3629 -> 4843
const FixedC2C<double, 2400, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 10> d2400;  // 3974 -> 4983
const FixedC2C<double, 2592, 216, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 6, 6, 6> d2592;  // 3671 -> 4942
const FixedC2C<double, 3000, 200, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 8, 5, 5> d3000;  // 3303 -> 4022
const FixedC2C<double, 3240, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6, 6> d3240;  // 2250 -> 4037
const FixedC2C<double, 3645, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 5> d3645;  // 2223 -> 3611
const FixedC2C<double, 4000, 400, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10, 4> d4000;  // 2561 -> 4242
const FixedC2C<double, 4800, 320, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 5, 4> d4800;  // 2457 -> 4206
const FixedC2C<double, 5184, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6, 6> d5184;  // 2728 -> 4497
const FixedC2C<double, 6144, 410, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 8, 3> d6144;  // 2746 -> 4153
const FixedC2C<double, 7200, 480, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15, 2> d7200;  // 2765 -> 3648
}  // namespace
}  // namespace sfft
