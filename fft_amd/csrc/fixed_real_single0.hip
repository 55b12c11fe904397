// fixed_real_single0.hip -- compile-time plans, real FP32 (half-length cores
// h, real length 2h), single pass: part 1 of 4 of the fast even lengths
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
const FixedReal<float, 135, 15, 17, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 9> r270;  // 3627 / 3331 -> 4170 / 3574
const FixedReal<float, 192, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12> r384;  // 4527 / 4594 -> 4831 / 4829
const FixedReal<float, 288, 24, 10, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 2> r576;  // 3901 / 3587 -> 4535 / 4106
const FixedReal<float, 384, 16, 16, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 12> r768;  // 3972 / 3672 -> 4269 / 4382
const FixedReal<float, 486, 54, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 9, 9, 6> r972;  // 3246 / 2957 -> 4340 / 3633
const FixedReal<float, 640, 43, 5, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 8, 5> r1280;  // 3814 / 3605 -> 4057 / 3915
const FixedReal<float, 768, 52, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 3> r1536;  // 3527 / 3309 -> 3778 / 3801
const FixedReal<float, 960, 64, 4, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 15, 4> r1920;  // 4055 / 4007 -> 4272 / 4201
const FixedReal<float, 1152, 96, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 6> r2304;  // 3807 / 3654 -> 4127 / 4331
const FixedReal<float, 1296, 144, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 12, 9> r2592;  // 2833 / 2742 -> 3649 / 3579
const FixedReal<float, 1536, 128, 2, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 6> r3072;  // 3786 / 3744 -> 4143 / 4109
const FixedReal<float, 2250, 225, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 10> r4500;  // 3173 / 2923 -> 3303 / 3281
const FixedReal<float, 2560, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 10> r5120;  // 3607 / 3492 -> 4100 / 4215
const FixedReal<float, 3072, 256, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 12> r6144;  // 3924 / 3792 -> 4305 / 4067
const FixedReal<float, 3888, 432, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 12, 9, 6, 6> r7776;  // 2360 / 2168 -> 3045 / 2210
const FixedReal<float, 4500, 300, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 5, 4> r9000;  // 2171 / 2055 -> 2374 / 2131
const FixedReal<float, 5120, 342, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 5, 4> r10240;  // 2501 / 2406 -> 2854 / 2526
const FixedReal<float, 6075, 405, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 15, 15, 3, 3, 3> r12150;  // 2324 / 2147 -> 2692 / 2183
const FixedReal<float, 6912, 576, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 12, 6, 6> r13824;  // 2020 / 2041 -> 3599 / 2720
const FixedReal<float, 8100, 900, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 9, 9> r16200;  // 2139 / 2008 -> 2603 / 2686
const FixedReal<float, 10000, 1000, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 10, 10, 10, 10> r20000;  // 2297 / 2063 -> 2603 / 2635
const FixedReal<float, 11520, 768, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 15, 3> r23040;  // 2308 / 2239 -> 3073 / 3034
const FixedReal<float, 12800, 512, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 32, 25, 16> r25600;  // 2158 / 2030 -> 2781 / 2921
const FixedReal<float, 16384, 1024, 1, Sched<0, 0, 0, 0, kLdsPadShift>, 0, 16, 16, 16, 4> r32768;  // 2556 / 2370 -> 2856 / 3171
}  // namespace
}  // namespace sfft
