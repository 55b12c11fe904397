// bluestein.hip -- lengths with prime factors the Stockham kernels do not
// plan (17, 19, 23, ..., and large primes), by the chirp-z identity
//   n k = (n^2 + k^2 - (k - n)^2) / 2
//   X[k] = w[k] * sum_n (x[n] w[n]) conj(w[k - n]),   w[n] = e^{-i pi n^2 / N},
// i.e. a cyclic convolution of length M >= 2N - 1 done with two complex FFTs
// of a planned length M (Bluestein, "A linear filtering approach to the
// computation of the discrete Fourier transform", 1970).  The reference
// instead runs a generic O(R^2) radix step for such factors
// (signalsmith-fft.h:187-215); the results agree to rounding.
#include "kernels.h"

namespace sfft {

// w[n] = e^{-i pi (n^2 mod 2N) / N} in FP64, rounded once (n^2 < 2^54).
template <typename T>
__global__ void k_bs_chirp(C2<T> *w, uint64_t n_len) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t n = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_len; n += stride) {
        const uint64_t q = (n * n) % (2 * n_len);
        double s, c;
        sincospi((double)q / (double)n_len, &s, &c);
        w[n] = {(T)c, (T)(-s)};
    }
}

// Convolution kernel b[m] = conj(w[m]) for m < N, conj(w[M - m]) for
// m > M - N, else 0.
template <typename T>
__global__ void k_bs_kernel(C2<T> *b, const C2<T> *w, uint64_t n_len, uint64_t m_len) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < m_len; m += stride) {
        C2<T> v = {(T)0, (T)0};
        if (m < n_len)
            v = cconj(w[m]);
        else if (m > m_len - n_len)
            v = cconj(w[m_len - m]);
        b[m] = v;
    }
}

template <typename T>
__global__ void k_bs_scale(C2<T> *z, uint64_t m_len, T s) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t m = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; m < m_len; m += stride)
        z[m] = {z[m].x * s, z[m].y * s};
}

// a[b][m] = load(x[b], m) * w[m] for m < N, 0 up to M.  The load applies the
// conjugation of the inverse, the half-bin rotation of a forward
// ModifiedRealFFT, or the real inverse pre-process of bin m.
template <typename T>
__global__ void k_bs_pre(const C2<T> *in, int64_t in_dist, C2<T> *a, uint64_t m_len,
                         uint64_t n_len, uint64_t batch, const C2<T> *w, int mode, int conj_in,
                         int half_shift, const C2<T> *tmi, const C2<T> *rot) {
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.y; b < batch; b += gridDim.y) {
        const C2<T> *x = in + (int64_t)b * in_dist;
        C2<T> *y = a + b * m_len;
        for (uint64_t m = t0; m < m_len; m += nt) {
            C2<T> v = {(T)0, (T)0};
            if (m < n_len) {
                if (mode == kModeC2RLoad) {
                    v = cconj(real_pre_bin(x, tmi, (int)m, (int)n_len, half_shift != 0));
                } else {
                    v = x[m];
                    if (conj_in) v.y = -v.y;
                    if (mode == kModeR2CLoad && half_shift) v = cmul(v, rot[m]);
                }
                v = cmul(v, w[m]);
            }
            y[m] = v;
        }
    }
}

// c[b][k] *= bhat[k] (bhat = FFT_M(b) / M).
template <typename T>
__global__ void k_bs_mul(C2<T> *c, const C2<T> *bhat, uint64_t m_len, uint64_t batch) {
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.y; b < batch; b += gridDim.y) {
        C2<T> *z = c + b * m_len;
        for (uint64_t k = t0; k < m_len; k += nt) z[k] = cmul(z[k], bhat[k]);
    }
}

// y[b][k] = store(c[b][k] * w[k]), k < N: conjugation of the inverse, and
// the half-bin rotation of an inverse ModifiedRealFFT.
template <typename T>
__global__ void k_bs_post(const C2<T> *c, uint64_t m_len, C2<T> *out, int64_t out_dist,
                          uint64_t n_len, uint64_t batch, const C2<T> *w, int conj_out,
                          const C2<T> *rotc) {
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t nt = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = blockIdx.y; b < batch; b += gridDim.y) {
        const C2<T> *z = c + b * m_len;
        C2<T> *y = out + (int64_t)b * out_dist;
        for (uint64_t k = t0; k < n_len; k += nt) {
            C2<T> v = cmul(z[k], w[k]);
            if (conj_out) v.y = -v.y;
            if (rotc) v = cmulc(v, rotc[k]);
            y[k] = v;
        }
    }
}

static dim3 grid_1d(uint64_t count) {
    return dim3((unsigned)std::min<uint64_t>((count + 255) / 256, 8192));
}
static dim3 grid_rows(uint64_t len, uint64_t batch) {
    const unsigned gx = (unsigned)std::min<uint64_t>((len + 255) / 256, 1024);
    const unsigned gy = (unsigned)std::min<uint64_t>(batch, 65535);
    return dim3(gx, gy);
}

template <typename T>
hipError_t bs_gen_chirp(C2<T> *w, C2<T> *b, size_t n_len, size_t m_len, hipStream_t stream) {
    hipLaunchKernelGGL((k_bs_chirp<T>), grid_1d(n_len), dim3(256), 0, stream, w, (uint64_t)n_len);
    hipLaunchKernelGGL((k_bs_kernel<T>), grid_1d(m_len), dim3(256), 0, stream, b, (const C2<T> *)w,
                       (uint64_t)n_len, (uint64_t)m_len);
    return hipGetLastError();
}

template <typename T>
hipError_t bs_scale(C2<T> *z, size_t m_len, double s, hipStream_t stream) {
    hipLaunchKernelGGL((k_bs_scale<T>), grid_1d(m_len), dim3(256), 0, stream, z, (uint64_t)m_len,
                       (T)s);
    return hipGetLastError();
}

template <typename T>
hipError_t bs_pre(const C2<T> *in, int64_t in_dist, C2<T> *a, size_t m_len, size_t n_len,
                  size_t batch, const C2<T> *w, int mode, int conj_in, int half_shift,
                  const C2<T> *tmi, const C2<T> *rot, hipStream_t stream) {
    hipLaunchKernelGGL((k_bs_pre<T>), grid_rows(m_len, batch), dim3(256), 0, stream, in, in_dist,
                       a, (uint64_t)m_len, (uint64_t)n_len, (uint64_t)batch, w, mode, conj_in,
                       half_shift, tmi, rot);
    return hipGetLastError();
}

template <typename T>
hipError_t bs_mul(C2<T> *c, const C2<T> *bhat, size_t m_len, size_t batch, hipStream_t stream) {
    hipLaunchKernelGGL((k_bs_mul<T>), grid_rows(m_len, batch), dim3(256), 0, stream, c, bhat,
                       (uint64_t)m_len, (uint64_t)batch);
    return hipGetLastError();
}

template <typename T>
hipError_t bs_post(const C2<T> *c, size_t m_len, C2<T> *out, int64_t out_dist, size_t n_len,
                   size_t batch, const C2<T> *w, int conj_out, const C2<T> *rotc,
                   hipStream_t stream) {
    hipLaunchKernelGGL((k_bs_post<T>), grid_rows(n_len, batch), dim3(256), 0, stream, c,
                       (uint64_t)m_len, out, out_dist, (uint64_t)n_len, (uint64_t)batch, w,
                       conj_out, rotc);
    return hipGetLastError();
}

#define SFFT_BS_INSTANTIATE(T)                                                                  \
    template hipError_t bs_gen_chirp<T>(C2<T> *, C2<T> *, size_t, size_t, hipStream_t);       \
    template hipError_t bs_scale<T>(C2<T> *, size_t, double, hipStream_t);                    \
    template hipError_t bs_pre<T>(const C2<T> *, int64_t, C2<T> *, size_t, size_t, size_t,    \
                                  const C2<T> *, int, int, int, const C2<T> *, const C2<T> *, \
                                  hipStream_t);                                               \
    template hipError_t bs_mul<T>(C2<T> *, const C2<T> *, size_t, size_t, hipStream_t);       \
    template hipError_t bs_post<T>(const C2<T> *, size_t, C2<T> *, int64_t, size_t, size_t,   \
                                   const C2<T> *, int, const C2<T> *, hipStream_t);
SFFT_BS_INSTANTIATE(float)
SFFT_BS_INSTANTIATE(double)

}  // namespace sfft
