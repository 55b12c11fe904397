// kernels.h -- host-side launch helpers for the device code in kernels.hip.
#pragma once

#include <hip/hip_runtime.h>

#include "stockham.h"

namespace sfft {

// Elements-per-thread instances compiled for the pass kernel.
constexpr int kEptSmall = 16;
constexpr int kEptLarge = 32;

// Block size limit of the pass kernel: 1024 threads for float; 512 for
// double, whose radix-16 butterflies and twiddles need more than 128
// registers (a single pass holds at most ~9600 double values in LDS, so no
// double plan needs more than 512 threads).
template <typename T>
constexpr int kMaxBlock() {
    return sizeof(T) == 8 ? 512 : 1024;
}

template <typename T>
hipError_t launch_pass(const PassArgs<T> &a, int ept, unsigned grid, unsigned block,
                       size_t lds_bytes, hipStream_t stream);

// 16-bit HBM storage with float arithmetic (half.hip); a.in / a.out address
// 4-byte complex values.
constexpr int kStorageF16 = 1;
constexpr int kStorageBF16 = 2;
// Block size limit of the 16-bit storage instances (room for the MFMA
// stage's operands without spilling).
constexpr int kMaxBlock16 = 512;
hipError_t launch_pass_16(const PassArgs<float> &a, int storage, int ept, unsigned grid,
                          unsigned block, size_t lds_bytes, hipStream_t stream);

// FP32 HBM storage with FP64 arithmetic (f32x.hip); a.in / a.out address
// 8-byte complex values, the descriptor's tables are FP64.
hipError_t launch_pass_f32x(const PassArgs<double> &a, unsigned grid, unsigned block,
                            size_t lds_bytes, hipStream_t stream);

// In-place forward real post-process of batch spectra of length h (distance
// dist complex values) after a two-pass complex FFT.
template <typename T>
hipError_t launch_real_post(C2<T> *out, int64_t dist, int h, size_t batch, const C2<T> *tmi,
                            int half_shift, hipStream_t stream);

// Stage twiddles W_{ns R}^{r k} for k < ns, r = 1..R-1 at dst[off + k tw_group(R) + r - 1],
// computed in FP64 (sincospi) and rounded once to T.
template <typename T>
hipError_t gen_stage_twiddles(C2<T> *dst, int off, int ns, int radix, hipStream_t stream);

// twiddlesMinusI[i] = -i e^{-2 pi i (i + s) / n}, i <= n/4, s = 1/2 for the
// half-shift kind (signalsmith-fft.h:420-425), and the half-shift
// rotations rot[i] = e^{-2 pi i * i / n}, i < n/2 (:426-432).
template <typename T>
hipError_t gen_real_twiddles(C2<T> *tmi, C2<T> *rot, size_t n, int half_shift,
                             hipStream_t stream);

// Two-level table for W_N^m: fine[m] for m < 2^shift,
// coarse[q] = W_N^{q 2^shift} for q < ceil(N / 2^shift); computed in FP64,
// rounded once to T.
template <typename T>
hipError_t gen_post_twiddles(C2<T> *fine, C2<T> *coarse, int shift, size_t N,
                             hipStream_t stream);
// Full table full[c n1 + e] = W_N^{c e} for c < N / n1, e < n1 (FP64, rounded
// once to T).
template <typename T>
hipError_t gen_post_twiddles_full(C2<T> *full, size_t n1, size_t N, hipStream_t stream);

// Bluestein (chirp-z) helpers, bluestein.hip.  w: chirp of length N;
// b: convolution kernel of length M; a / c: per-transform rows of length M.
template <typename T>
hipError_t bs_gen_chirp(C2<T> *w, C2<T> *b, size_t n_len, size_t m_len, hipStream_t stream);
template <typename T>
hipError_t bs_scale(C2<T> *z, size_t m_len, double s, hipStream_t stream);
template <typename T>
hipError_t bs_pre(const C2<T> *in, int64_t in_dist, C2<T> *a, size_t m_len, size_t n_len,
                  size_t batch, const C2<T> *w, int mode, int conj_in, int half_shift,
                  const C2<T> *tmi, const C2<T> *rot, hipStream_t stream);
template <typename T>
hipError_t bs_mul(C2<T> *c, const C2<T> *bhat, size_t m_len, size_t batch, hipStream_t stream);
template <typename T>
hipError_t bs_post(const C2<T> *c, size_t m_len, C2<T> *out, int64_t out_dist, size_t n_len,
                   size_t batch, const C2<T> *w, int conj_out, const C2<T> *rotc,
                   hipStream_t stream);

// Counter-based splitmix64 inputs (same stream as fft_amd/synth.py).
hipError_t fill_synthetic(void *dst, size_t count, int f64, int dist, uint64_t seed,
                          hipStream_t stream);
// The same for n buffers in one launch; `entries` is device memory, and
// max_count bounds their counts (it sizes the grid; f64 as for the entries).
struct FillEntry {
    void *dst;
    uint64_t count;  // values of the real type
    uint64_t seed;
    int32_t f64, dist;
};
hipError_t fill_synthetic_many(const FillEntry *entries, int n, uint64_t max_count, int f64,
                               hipStream_t stream);

}  // namespace sfft
