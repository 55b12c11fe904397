// fixed_col_mixed.hip -- compile-time column passes of two-pass plans for the
// mixed-radix column lengths the planner picks (by bytes) for complex FP32
// lengths between 16384 and 2^18; radices as the runtime planner's cost
// model chooses them at 16 elements per thread.  Unpadded LDS: over the 185
// fast lengths in (16384, 2^18] two-pass throughput is 1730 vs 1719 GB/s with
// the one-in-16 pad (up to 8 % faster on single lengths).  Columns per
// block: the widest tile that keeps two blocks per CU measured fastest for
// most lengths (324, 432, 729 at twice the former width: pass times -20 to
// -32 %), except with radix-25 stages (250 at 32 columns instead of 16:
// 1.5-1.8x slower, fewer resident waves) and where it was neutral (243).
// The first pass reads its input non-temporally (Sched NT 1): two-pass
// lengths in [16000, 2^18] +6 % (median), R2C +4 %.
#include "fixed.h"

namespace sfft {
namespace {
const FixedCol<float, 480, 32, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 15, 2> col480;
const FixedCol<float, 375, 25, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 15> col375;
const FixedCol<float, 486, 54, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9, 6> col486;
const FixedCol<float, 405, 45, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9, 5> col405;
const FixedCol<float, 500, 50, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 10, 10, 5> col500;
const FixedCol<float, 400, 25, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 16> col400;
const FixedCol<float, 384, 26, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 8, 3> col384;
const FixedCol<float, 432, 36, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 12, 6, 6> col432;
const FixedCol<float, 450, 30, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 15, 2> col450;
const FixedCol<float, 324, 36, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 6, 6> col324;
const FixedCol<float, 320, 22, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 5, 4> col320;
const FixedCol<float, 250, 25, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 10> col250;
const FixedCol<float, 243, 27, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9, 3> col243;
const FixedCol<float, 360, 24, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 8, 3> col360;
const FixedCol<float, 300, 20, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 5, 4> col300;
const FixedCol<float, 270, 23, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 6, 3> col270;
const FixedCol<float, 288, 24, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 12, 12, 2> col288;
const FixedCol<float, 625, 25, 8, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 25> col625;
const FixedCol<float, 225, 15, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 15> col225;
const FixedCol<float, 240, 16, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 15> col240;
const FixedCol<float, 200, 13, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 8> col200;
const FixedCol<float, 216, 18, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 6, 6, 6> col216;
const FixedCol<float, 192, 16, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 12> col192;
const FixedCol<float, 180, 15, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 12> col180;
const FixedCol<float, 162, 18, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9, 2> col162;
const FixedCol<float, 160, 16, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 16, 10> col160;
const FixedCol<float, 150, 15, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 10> col150;
const FixedCol<float, 540, 45, 16, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 6, 6> col540;
// 125: five threads per column (one radix-25 butterfly each, then five
// radix-5 ones) and 64 columns a block: two-pass lengths with a 125 pass
// 4-10 % faster than nine threads per column at 32 columns (16875: 2.04 ->
// 2.23 TB/s, 31250: 2.03 -> 2.21), same bits.  The same change for other
// lengths with idle threads in a stage measured slower (150 at five threads
// and 64 columns, 135 at nine, 250 at ten, 324 at 18: medians -2 to -9 %),
// and 64 columns a block for 81, 135, 144, 150 and 160 was neutral (medians
// 0.98-1.01).
const FixedCol<float, 125, 5, 64, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 25, 5> col125;
const FixedCol<float, 729, 81, 8, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9, 9> col729;
const FixedCol<float, 144, 12, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 12, 12> col144;
const FixedCol<float, 135, 15, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 15, 9> col135;
const FixedCol<float, 81, 9, 32, Sched<0, 0, 0, 0, kLdsNoPad, 1>, 0, 9, 9> col81;
}  // namespace
}  // namespace sfft
