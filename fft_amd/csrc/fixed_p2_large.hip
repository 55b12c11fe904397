// fixed_p2_large.hip -- compile-time plans, complex FP32: 16384, 8192.
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 16384, 1024, 1, 16, 16, 16, 4> c16384;
const FixedC2C<float, 8192, 512, 1, 16, 16, 8, 4> c8192;
}  // namespace
}  // namespace sfft
