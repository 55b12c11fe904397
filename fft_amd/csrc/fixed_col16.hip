// fixed_col16.hip -- compile-time column passes of two-pass plans with 16-bit
// storage: IEEE half or bfloat16 pairs in HBM, float butterflies and float
// twiddles (the same arithmetic as the runtime-planned 16-bit pass, rounded to
// the storage format once per pass).  Same shapes as the FP32 column kernels;
// column length 2048 keeps the runtime-planned pass, which measured faster
// with 16-bit storage (0.154 vs 0.161 ms at 2^22 x 4).
#include "fixed.h"

namespace sfft {
namespace {
template <class S>
struct Col16 {
    using E = Stored<S>;
    const FixedCol<E, 128, 8, 32, Sched<0, 0, 0, 0>, 0, 16, 8> c128;
    const FixedCol<E, 256, 16, 16, Sched<0, 0, 0, 0>, 0, 16, 16> c256;
    const FixedCol<E, 512, 32, 16, Sched<0, 0, 0, 0>, 0, 16, 16, 2> c512;
    const FixedCol<E, 1024, 64, 8, Sched<0, 0, 0, 0>, 0, 16, 16, 4> c1024;
};
const Col16<StoreF16> col16_f16;
const Col16<StoreBF16> col16_bf16;
}  // namespace
}  // namespace sfft
