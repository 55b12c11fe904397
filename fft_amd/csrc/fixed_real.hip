// fixed_real.hip -- compile-time plans, real FP32 (half-length cores): 2048 (1024 = 16 16 4).
#include "fixed.h"

namespace sfft {
namespace {
const FixedReal<float, 1024, 64, 4, 16, 16, 4> r2048;
}  // namespace
}  // namespace sfft
