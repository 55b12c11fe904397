// fixed_real_single3.hip -- compile-time plans, real FP32 (half-length cores
// h, real length 2h), single pass: part 4 of 4 of the fast even lengths
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
const FixedReal<float, 162, 18, 14, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 2> r324;  // 3334 / 3037 -> 4454 / 3610
const FixedReal<float, 270, 23, 11, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 3> r540;  // 3620 / 3129 -> 3939 / 3393
const FixedReal<float, 324, 36, 7, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6> r648;  // 3594 / 3066 -> 4455 / 3616
const FixedReal<float, 480, 32, 8, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 2> r960;  // 4065 / 3949 -> 4317 / 4235
const FixedReal<float, 576, 48, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6, 6> r1152;  // 3803 / 3617 -> 4524 / 4292
const FixedReal<float, 729, 81, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9> r1458;  // 3660 / 3214 -> 4453 / 4011
const FixedReal<float, 864, 72, 3, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6> r1728;  // 3649 / 3324 -> 4135 / 4001
const FixedReal<float, 1080, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 12, 6> r2160;  // 3539 / 3290 -> 3824 / 3530
const FixedReal<float, 1280, 86, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 5> r2560;  // 3747 / 3598 -> 4090 / 4138
const FixedReal<float, 1500, 150, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 10, 10> r3000;  // 2983 / 2824 -> 3544 / 3104
const FixedReal<float, 2187, 243, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 3> r4374;  // 2987 / 2754 -> 4027 / 3726
const FixedReal<float, 2430, 270, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 3> r4860;  // 1683 / 1608 -> 2404 / 1799
const FixedReal<float, 2916, 324, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6, 6> r5832;  // 1960 / 1817 -> 2706 / 1919
const FixedReal<float, 3645, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 5> r7290;  // 2375 / 2138 -> 3212 / 2463
const FixedReal<float, 4320, 360, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 6, 5> r8640;  // 2360 / 2253 -> 3095 / 2478
const FixedReal<float, 4860, 540, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9, 6> r9720;  // 1840 / 1638 -> 3319 / 2405
const FixedReal<float, 5832, 648, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 9, 6> r11664;  // 2021 / 1849 -> 3550 / 2513
const FixedReal<float, 6561, 729, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 9> r13122;  // 2195 / 2114 -> 3695 / 3037
const FixedReal<float, 8000, 320, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 10> r16000;  // 2092 / 1959 -> 2676 / 2559
const FixedReal<float, 9600, 640, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 8, 5> r19200;  // 2129 / 2147 -> 3079 / 2953
const FixedReal<float, 10800, 720, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15, 3> r21600;  // 2206 / 2127 -> 3068 / 2891
const FixedReal<float, 12288, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 32, 12> r24576;  // 2338 / 2215 -> 2885 / 2911
const FixedReal<float, 15360, 1024, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 15, 4> r30720;  // 2532 / 2395 -> 2889 / 3106
}  // namespace
}  // namespace sfft
