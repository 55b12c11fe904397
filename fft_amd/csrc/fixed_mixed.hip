// fixed_mixed.hip -- compile-time plans, complex FP32: 1920 = 16 12 10, 3000 = 12 10 25.
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 1920, 160, 2, 16, 12, 10> c1920;
const FixedC2C<float, 3000, 250, 1, 12, 10, 25> c3000;
}  // namespace
}  // namespace sfft
