// fixed_f64_single3.hip -- compile-time plans, complex FP64, single pass:
// part 4 of 4 of the fast lengths 2^a 3^b 5^c in [64, 10000] that have no
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
const FixedC2C<double, 81, 9, 56, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9> d81;  // 4392 -> 5213
const FixedC2C<double, 108, 12, 42, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9> d108;  // 4484 -> 5226
const FixedC2C<double, 150, 15, 17, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10> d150;  // 4623 -> 5198
const FixedC2C<double, 192, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12> d192;  // 4973 -> 5305
const FixedC2C<double, 243, 27, 18, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 3> d243;  // 3565 -> 5181
const FixedC2C<double, 300, 20, 12, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 5, 4> d300;  // 4016 -> 5056
const FixedC2C<double, 375, 25, 12, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 5, 5> d375;  // 3944 -> 4902
const FixedC2C<double, 432, 36, 11, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 6, 6> d432;  // 4441 -> 5214
const FixedC2C<double, 500, 50, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 5> d500;  // 4083 -> 5152
const FixedC2C<double, 625, 42, 6, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 5, 5, 5, 5> d625;  // 3724 -> 4090
const FixedC2C<double, 720, 48, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 3> d720;  // 4556 -> 5091
const FixedC2C<double, 800, 80, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 8> d800;  // 3878 -> 5136
const FixedC2C<double, 960, 64, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 4> d960;  // 4528 -> 5079
const FixedC2C<double, 1125, 75, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 5> d1125;  // 4151 -> 4784
const FixedC2C<double, 1250, 84, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 5, 5, 5, 5, 2> d1250;  // 3347 -> 3680
const FixedC2C<double, 1440, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 10> d1440;  // 3218 -> 5040
const FixedC2C<double, 1600, 160, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 10, 10> d1600;  // 3395 -> 5150
const FixedC2C<double, 1875, 125, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 5, 5, 5> d1875;  // 3580 -> 4573
const FixedC2C<double, 2025, 135, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 3, 3> d2025;  // 2795 -> 3300
const FixedC2C<double, 2304, 192, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 12> d2304;  // 3989 -> 5091
const FixedC2C<double, 2560, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 10> d2560;  // 4146 -> 5165
const FixedC2C<double, 2916, 324, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6, 6> d2916;  // 2214 -> 3788
const FixedC2C<double, 3200, 214, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 8, 5, 5> d3200;  // 3672 -> 4547
const FixedC2C<double, 3600, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15> d3600;  // 4260 -> 4799
const FixedC2C<double, 3888, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6, 6> d3888;  // 2489 -> 4336
const FixedC2C<double, 4608, 384, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 12, 2> d4608;  // 2459 -> 4268
const FixedC2C<double, 5120, 342, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 5, 4> d5120;  // 2624 -> 4145
const FixedC2C<double, 6000, 400, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 5, 5> d6000;  // 2707 -> 4051
const FixedC2C<double, 6912, 461, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 3, 3, 3> d6912;  // 2557 -> 2964
}  // namespace
}  // namespace sfft
