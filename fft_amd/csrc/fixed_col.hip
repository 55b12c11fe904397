// fixed_col.hip -- compile-time column passes of two-pass plans, complex
// FP32 and FP64.  FixedCol<type, column length N, threads per column, columns per
// block, Sched<...>, variant, radices...>; a two-pass plan whose column
// length has one uses it for that pass (same radices as the runtime-planned
// pass, so the same bits).  Column length 4096 keeps the runtime-planned
// pass: a compiled 4096 column kernel (tiles of 4 columns) measured slower
// (2^24: 1070 -> 921 GB/s).  Tile widths: 2048-length columns 8 wide and
// 256-length columns 32 wide measured faster than half that (2^22 +15 %,
// 2^21 +6 %, 2^15..2^16 +2-4 %); 512 and 1024 wider measured slower or
// the same.  FP32 lengths 128-512 read the first pass' input non-temporally
// (Sched NT 1: 2^19 x 64 and 2^18 x 128 +13-15 %); for 1024 and 2048 that
// measured 3-4 % slower (2^20, 2^22), and non-temporal output stores of the
// second pass slowed the lengths whose output is read back (the real
// post-process) and some others (up to 8 %).
#include "fixed.h"

namespace sfft {
namespace {
const FixedCol<float, 128, 8, 32, Sched<0, 0, 0, 0, kLdsPadShift, 1>, 0, 16, 8> col128;
const FixedCol<float, 256, 16, 32, Sched<0, 0, 0, 0, kLdsPadShift, 1>, 0, 16, 16> col256;
const FixedCol<float, 512, 32, 16, Sched<0, 0, 0, 0, kLdsPadShift, 1>, 0, 16, 16, 2> col512;
const FixedCol<float, 1024, 64, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 4> col1024;
const FixedCol<float, 2048, 128, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 8> col2048;
// complex FP64 (16-byte values: half the columns per LDS budget), first-pass
// input loads non-temporal (2^20 x 16 1.75 -> 2.02 TB/s, 2^18 x 64 2.22 ->
// 2.50); length 256 keeps the runtime-planned pass (compiled: 2^16 x 256
// 2.47 -> 2.10 TB/s, 2^17 x 128 2.22 -> 2.04)
const FixedCol<double, 512, 32, 8, Sched<0, 0, 0, 0, kLdsPadShift, 1>, 0, 16, 16, 2> col512d;
const FixedCol<double, 1024, 64, 4, Sched<0, 0, 0, 0, kLdsPadShift, 1>, 0, 16, 16, 4> col1024d;
}  // namespace
}  // namespace sfft
