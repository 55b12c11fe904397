// fixed_real_single1.hip -- compile-time plans, real FP32 (half-length cores
// h, real length 2h), single pass: part 2 of 4 of the fast even lengths
// in [270, 37500] that have no hand-tuned kernel.  For each length two
// candidates were measured against the runtime-planned pass (batched R2C
// forward and C2R inverse launches of about 256 MB, cold inputs): the radices
// the runtime cost model picks for h at 16 elements per thread and at 32
// (radix 32 allowed), threads per transform as that model computes them, and
// about 256 threads per block.  A length is listed only where its faster
// candidate beat the runtime pass by 5 % or more over both directions, with
// neither direction 3 % slower (forward / inverse GB/s, runtime -> compiled,
// follow each line).  The other lengths keep the runtime pass (their
// candidates lost, up to 2.5x slower above h = 10000).
// FixedReal<type, h, threads per transform, transforms per block,
//           Sched<prefetch stage, min waves per SIMD, pair loads, prefetch first,
//           LDS pad shift>, variant, radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedReal<float, 150, 15, 17, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10> r300;  // 3754 / 3384 -> 4106 / 3460
const FixedReal<float, 216, 18, 14, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 6, 6, 6> r432;  // 3977 / 3642 -> 4448 / 3852
const FixedReal<float, 300, 13, 19, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 25, 12> r600;  // 3789 / 3346 -> 4181 / 3902
const FixedReal<float, 405, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 5> r810;  // 3419 / 3079 -> 4603 / 3859
const FixedReal<float, 500, 50, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 5> r1000;  // 3898 / 3338 -> 4652 / 4090
const FixedReal<float, 648, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6> r1296;  // 3120 / 2982 -> 4063 / 3574
const FixedReal<float, 800, 32, 8, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25> r1600;  // 3726 / 3622 -> 4075 / 4036
const FixedReal<float, 972, 108, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9> r1944;  // 3346 / 3094 -> 4109 / 3900
const FixedReal<float, 1200, 80, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 5> r2400;  // 3898 / 3830 -> 4246 / 4166
const FixedReal<float, 1440, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 10> r2880;  // 3020 / 2915 -> 3907 / 3810
const FixedReal<float, 1728, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12> r3456;  // 3205 / 3185 -> 3797 / 3846
const FixedReal<float, 2304, 192, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 12> r4608;  // 3842 / 3858 -> 4298 / 4315
const FixedReal<float, 2592, 216, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 6, 6, 6> r5184;  // 3072 / 3010 -> 3769 / 3376
const FixedReal<float, 3240, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6, 6> r6480;  // 1793 / 1752 -> 2283 / 1737
const FixedReal<float, 4000, 400, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10, 4> r8000;  // 2466 / 2210 -> 3027 / 2314
const FixedReal<float, 4608, 192, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 12, 12> r9216;  // 2403 / 2364 -> 3494 / 3378
const FixedReal<float, 5184, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6, 6> r10368;  // 2694 / 2539 -> 3305 / 2636
const FixedReal<float, 6144, 410, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 8, 3> r12288;  // 2795 / 2606 -> 3150 / 2901
const FixedReal<float, 7290, 810, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 9> r14580;  // 2186 / 2060 -> 2634 / 2817
const FixedReal<float, 8640, 720, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12, 5> r17280;  // 2111 / 2009 -> 3188 / 2852
const FixedReal<float, 10240, 683, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 8, 5> r20480;  // 2213 / 2131 -> 2986 / 2997
const FixedReal<float, 11664, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 9, 9, 9> r23328;  // 1800 / 1716 -> 2583 / 2717
const FixedReal<float, 12960, 480, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 9, 9, 5> r25920;  // 1826 / 1719 -> 2519 / 2729
}  // namespace
}  // namespace sfft
