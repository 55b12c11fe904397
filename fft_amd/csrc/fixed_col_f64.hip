// fixed_col_f64.hip -- compile-time column passes of two-pass complex FP64
// plans for the column lengths 512 and 1024: the radices of the FP32 column
// kernels (radix 16 at most, as the FP64 runtime planner allows), half as
// many columns per block (the same LDS bytes per block).  Measured against
// the runtime-planned columns (bitwise equal): 2^18 x 64 1.77 -> 2.27 TB/s,
// 2^20 x 16 1.65 -> 1.72.  Lengths 128, 256 and 2048 were slower compiled
// (2^16 x 256 2.49 -> 2.10 at 8 or 16 columns, 2^22 x 4 1.49 -> 1.33-1.44 at
// 2 or 4 columns) and keep the runtime-planned pass, as do passes with fewer
// than 512 columns per transform (the planner's rule).
// FixedCol<type, N, threads per column, columns per block, Sched<...>, variant, radices...>
#include "fixed.h"

namespace sfft {
namespace {
const FixedCol<double, 512, 32, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 2> dcol512;
const FixedCol<double, 1024, 64, 4, Sched<0, 0, 0, 0>, 0, 16, 16, 4> dcol1024;
}  // namespace
}  // namespace sfft
