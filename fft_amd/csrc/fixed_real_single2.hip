// fixed_real_single2.hip -- compile-time plans, real FP32 (half-length cores
// h, real length 2h), single pass: part 3 of 4 of the fast even lengths
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
const FixedReal<float, 160, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 10> r320;  // 4169 / 4063 -> 4600 / 4502
const FixedReal<float, 243, 27, 9, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 3> r486;  // 3612 / 3281 -> 4799 / 3815
const FixedReal<float, 320, 22, 11, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 5, 4> r640;  // 3975 / 3621 -> 4101 / 3898
const FixedReal<float, 432, 36, 7, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 6, 6> r864;  // 4134 / 3698 -> 4478 / 4082
const FixedReal<float, 540, 45, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 6, 6> r1080;  // 3417 / 3106 -> 3741 / 3169
const FixedReal<float, 720, 48, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 3> r1440;  // 3962 / 3793 -> 4292 / 4153
const FixedReal<float, 810, 90, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 9, 9> r1620;  // 3435 / 3283 -> 4363 / 4047
const FixedReal<float, 1000, 100, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10> r2000;  // 3304 / 2966 -> 4130 / 3695
const FixedReal<float, 1215, 135, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 9, 9> r2430;  // 2602 / 2541 -> 3145 / 2971
const FixedReal<float, 1458, 162, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 9, 2> r2916;  // 2531 / 2465 -> 3615 / 3152
const FixedReal<float, 1944, 216, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 6, 6, 6> r3888;  // 2618 / 2530 -> 3651 / 2978
const FixedReal<float, 2400, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 10> r4800;  // 3451 / 3383 -> 3910 / 3756
const FixedReal<float, 2880, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 12> r5760;  // 3835 / 3577 -> 3961 / 3912
const FixedReal<float, 3456, 288, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 6, 6, 6> r6912;  // 1987 / 1984 -> 2744 / 2158
const FixedReal<float, 4050, 338, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 6, 3> r8100;  // 2104 / 1903 -> 2337 / 1970
const FixedReal<float, 4800, 320, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 5, 4> r9600;  // 2407 / 2415 -> 2860 / 2648
const FixedReal<float, 5760, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 15, 12> r11520;  // 2649 / 2591 -> 2673 / 3576
const FixedReal<float, 6480, 240, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 9, 9, 5> r12960;  // 1824 / 1710 -> 2473 / 3562
const FixedReal<float, 7776, 288, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 9, 9, 3> r15552;  // 2138 / 2079 -> 2509 / 2485
const FixedReal<float, 9216, 768, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 6, 6> r18432;  // 2053 / 2087 -> 3208 / 2832
const FixedReal<float, 10368, 864, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 12, 6> r20736;  // 2279 / 2168 -> 2439 / 2811
const FixedReal<float, 12000, 480, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 15> r24000;  // 2251 / 2111 -> 2724 / 3042
const FixedReal<float, 14400, 960, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 15, 4> r28800;  // 2364 / 2262 -> 2799 / 2339
}  // namespace
}  // namespace sfft
