// fixed_p2_mid.hip -- compile-time plans, complex FP32: 4096, 2048, 1024, 512, 256.
#include "fixed.h"

namespace sfft {
namespace {
const FixedC2C<float, 4096, 256, 1, 16, 16, 16> c4096;
const FixedC2C<float, 2048, 128, 2, 16, 16, 8> c2048;
const FixedC2C<float, 1024, 64, 4, 16, 16, 4> c1024;
const FixedC2C<float, 512, 32, 8, 16, 8, 4> c512;
const FixedC2C<float, 256, 16, 16, 16, 16> c256;
}  // namespace
}  // namespace sfft
