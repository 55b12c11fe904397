// fixed_col.hip -- compile-time column passes of two-pass plans, complex
// FP32 and FP64.  FixedCol<type, column length N, threads per column, columns per
// block, Sched<...>, variant, radices...>; a two-pass plan whose column
// length has one uses it for that pass (same radices as the runtime-planned
// pass, so the same bits).  Column length 4096 keeps the runtime-planned
// pass: its tiles of 2-4 columns measured faster there.
#include "fixed.h"

namespace sfft {
namespace {
const FixedCol<float, 128, 8, 32, Sched<0, 0, 0, 0>, 0, 16, 8> col128;
const FixedCol<float, 256, 16, 16, Sched<0, 0, 0, 0>, 0, 16, 16> col256;
const FixedCol<float, 512, 32, 16, Sched<0, 0, 0, 0>, 0, 16, 16, 2> col512;
const FixedCol<float, 1024, 64, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 4> col1024;
const FixedCol<float, 2048, 128, 4, Sched<0, 0, 0, 0>, 0, 16, 16, 8> col2048;
// complex FP64 (16-byte values: half the columns per LDS budget)
const FixedCol<double, 256, 16, 8, Sched<0, 0, 0, 0>, 0, 16, 16> col256d;
const FixedCol<double, 512, 32, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 2> col512d;
const FixedCol<double, 1024, 64, 4, Sched<0, 0, 0, 0>, 0, 16, 16, 4> col1024d;
}  // namespace
}  // namespace sfft
