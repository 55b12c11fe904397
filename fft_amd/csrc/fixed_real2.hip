// fixed_real2.hip -- compile-time plans, real FP32 lengths 1024 and 4096...16384
// (half-length complex cores with the shapes of the complex kernels; 32768
// is left to the runtime-planned pass: its forward core spills at any shape
// tried).
#include "fixed.h"

namespace sfft {
namespace {
const FixedReal<float, 512, 32, 8, Sched<0, 0, 0>, 0, 16, 8, 4> r1024a;
const FixedReal<float, 2048, 128, 2, Sched<-1, 0, 0>, 0, 16, 16, 8> r4096a;
const FixedReal<float, 4096, 256, 1, Sched<-1, 0, 0>, 0, 16, 16, 16> r8192a;
const FixedReal<float, 8192, 512, 1, Sched<-1, 0, 0>, 0, 16, 16, 8, 4> r16384a;
}  // namespace
}  // namespace sfft
