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

// single-pass complex plans with 16-bit storage (shapes of the FP32 kernels;
// 1.7-3.3x the runtime-planned 16-bit pass on 0.5 GiB batches)
template <class S>
struct One16 {
    using E = Stored<S>;
    const FixedC2C<E, 1024, 64, 4, Sched<0, 4, 0>, 0, 16, 16, 4> c1024;
    const FixedC2C<E, 4096, 256, 1, Sched<-1, 4, 0>, 0, 16, 16, 16> c4096;
    const FixedC2C<E, 16384, 1024, 1, Sched<0, 0, 0, 1>, 0, 16, 16, 16, 4> c16384;
    const FixedC2C<E, 256, 16, 16, Sched<-1, 0, 0>, 0, 16, 16> c256;
    const FixedC2C<E, 512, 32, 8, Sched<0, 4, 0>, 0, 16, 8, 4> c512;
    const FixedC2C<E, 2048, 128, 2, Sched<-1, 0, 0>, 0, 16, 16, 8> c2048;
    const FixedC2C<E, 8192, 512, 1, Sched<-1, 0, 0>, 0, 16, 16, 8, 4> c8192;
};
const One16<StoreF16> one16_f16;
const One16<StoreBF16> one16_bf16;
}  // namespace
}  // namespace sfft
