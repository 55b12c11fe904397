// kernels.hip -- pass kernel, FP64 table generators and the synthetic input
// generator for gfx950.
#include "kernels.h"

namespace sfft {

// ---------------------------------------------------------------- passes --

template <typename T, int EPT>
__global__ void __launch_bounds__(kMaxBlock<T>()) k_pass(PassArgs<T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
    Block<T, EPT> blk(a, a.in, a.out, a.total_slots, lds, (int)threadIdx.x,
                      (int)xcd_tile(blockIdx.x, gridDim.x) * a.pd.tpw);
    blk.run();
}

template <typename T, int EPT>
static hipError_t launch_one(const PassArgs<T> &a, unsigned grid, unsigned block, size_t lds_bytes,
                             hipStream_t stream) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_pass<T, EPT>) != hipSuccess) return lds_err;
    hipLaunchKernelGGL((k_pass<T, EPT>), dim3(grid), dim3(block), lds_bytes, stream, a);
    return hipGetLastError();
}

template <>
hipError_t launch_pass<float>(const PassArgs<float> &a, int ept, unsigned grid, unsigned block,
                              size_t lds_bytes, hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    if (ept <= kEptSmall) return launch_one<float, kEptSmall>(a, grid, block, lds_bytes, stream);
    return launch_one<float, kEptLarge>(a, grid, block, lds_bytes, stream);
}

// FP64 has only the 16-element instance (32 elements of double2 spill).
template <>
hipError_t launch_pass<double>(const PassArgs<double> &a, int ept, unsigned grid, unsigned block,
                               size_t lds_bytes, hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    if (ept > kEptSmall) return hipErrorInvalidValue;
    return launch_one<double, kEptSmall>(a, grid, block, lds_bytes, stream);
}


// ------------------------------------------- real post-process (two-pass) --

template <typename T>
__global__ void k_real_post(C2<T> *out, int64_t dist, int h, int batch, const C2<T> *tmi,
                            int mod) {
    const int t0 = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int nt = (int)(gridDim.x * blockDim.x);
    for (int b = (int)blockIdx.y; b < batch; b += (int)gridDim.y)
        real_post_inplace(out + (int64_t)b * dist, tmi, h, mod != 0, t0, nt);
}

template <typename T>
hipError_t launch_real_post(C2<T> *out, int64_t dist, int h, size_t batch, const C2<T> *tmi,
                            int half_shift, hipStream_t stream) {
    if (batch == 0) return hipSuccess;
    // about 2048 workgroups in all (at least 64 per transform), so that a
    // small batch of long transforms still covers every CU
    const int pairs = h / 2 + 1;
    const size_t want = batch >= 32 ? 64 : 2048 / batch;
    unsigned gx = (unsigned)((pairs + 255) / 256);
    if (gx > want) gx = (unsigned)want;
    const unsigned gy = (unsigned)(batch < 65535 ? batch : 65535);
    hipLaunchKernelGGL((k_real_post<T>), dim3(gx, gy), dim3(256), 0, stream, out, dist, h,
                       (int)batch, tmi, half_shift);
    return hipGetLastError();
}
template hipError_t launch_real_post<float>(C2<float> *, int64_t, int, size_t, const C2<float> *,
                                            int, hipStream_t);
template hipError_t launch_real_post<double>(C2<double> *, int64_t, int, size_t,
                                             const C2<double> *, int, hipStream_t);

// --------------------------------------------------------------- twiddles --

// e^{-2 pi i m / L} for an exact integer ratio, in FP64.
__device__ __forceinline__ C2<double> unit_root(uint64_t m, uint64_t L) {
    m %= L;
    double s, c;
    sincospi(-2.0 * (double)m / (double)L, &s, &c);
    return {c, s};
}

template <typename T>
__global__ void k_stage_tw(C2<T> *dst, int off, int ns, int radix) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int cnt = ns * (radix - 1);
    if (idx >= cnt) return;
    const int k = idx / (radix - 1);
    const int r = idx % (radix - 1) + 1;
    const C2<double> w = unit_root((uint64_t)r * (uint64_t)k, (uint64_t)ns * (uint64_t)radix);
    dst[off + k * tw_group(radix) + (r - 1)] = {(T)w.x, (T)w.y};
}

template <typename T>
hipError_t gen_stage_twiddles(C2<T> *dst, int off, int ns, int radix, hipStream_t stream) {
    const int cnt = ns * (radix - 1);
    if (cnt <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_stage_tw<T>), dim3((cnt + 255) / 256), dim3(256), 0, stream, dst, off,
                       ns, radix);
    return hipGetLastError();
}
template hipError_t gen_stage_twiddles<float>(C2<float> *, int, int, int, hipStream_t);
template hipError_t gen_stage_twiddles<double>(C2<double> *, int, int, int, hipStream_t);

template <typename T>
__global__ void k_real_tw(C2<T> *tmi, C2<T> *rot, uint64_t n, int half_shift) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i <= n / 4) {
        // -i e^{-2 pi i (i + s)/n} with s = 1/2: exponent over 2n is (2i + 1).
        const C2<double> w = half_shift ? unit_root(2 * i + 1, 2 * n) : unit_root(i, n);
        tmi[i] = {(T)w.y, (T)(-w.x)};
    }
    if (half_shift && i < n / 2) {
        const C2<double> w = unit_root(i, n);
        rot[i] = {(T)w.x, (T)w.y};
    }
}

template <typename T>
hipError_t gen_real_twiddles(C2<T> *tmi, C2<T> *rot, size_t n, int half_shift,
                             hipStream_t stream) {
    const size_t cnt = n / 2 + 1;
    hipLaunchKernelGGL((k_real_tw<T>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream,
                       tmi, rot, (uint64_t)n, half_shift);
    return hipGetLastError();
}
template hipError_t gen_real_twiddles<float>(C2<float> *, C2<float> *, size_t, int, hipStream_t);
template hipError_t gen_real_twiddles<double>(C2<double> *, C2<double> *, size_t, int,
                                              hipStream_t);

template <typename T>
__global__ void k_post_tw(C2<T> *fine, C2<T> *coarse, int shift, uint64_t N) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t nf = 1ull << shift;
    const uint64_t nc = (N + nf - 1) >> shift;
    if (i < nf) {
        const C2<double> w = unit_root(i, N);
        fine[i] = {(T)w.x, (T)w.y};
    }
    if (i < nc) {
        const C2<double> w = unit_root(i << shift, N);
        coarse[i] = {(T)w.x, (T)w.y};
    }
}

template <typename T>
hipError_t gen_post_twiddles(C2<T> *fine, C2<T> *coarse, int shift, size_t N,
                             hipStream_t stream) {
    const size_t nf = 1ull << shift;
    const size_t nc = (N + nf - 1) >> shift;
    const size_t cnt = nf > nc ? nf : nc;
    hipLaunchKernelGGL((k_post_tw<T>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, stream,
                       fine, coarse, shift, (uint64_t)N);
    return hipGetLastError();
}
template <typename T>
__global__ void k_post_tw_full(C2<T> *full, uint32_t n1, uint64_t N) {
    const uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (i < N) {
        const uint64_t c = i / n1, e = i - c * n1;
        const C2<double> w = unit_root(c * e, N);
        full[i] = {(T)w.x, (T)w.y};
    }
}

template <typename T>
hipError_t gen_post_twiddles_full(C2<T> *full, size_t n1, size_t N, hipStream_t stream) {
    hipLaunchKernelGGL((k_post_tw_full<T>), dim3((unsigned)((N + 255) / 256)), dim3(256), 0, stream,
                       full, (uint32_t)n1, (uint64_t)N);
    return hipGetLastError();
}
template hipError_t gen_post_twiddles_full<float>(C2<float> *, size_t, size_t, hipStream_t);
template hipError_t gen_post_twiddles_full<double>(C2<double> *, size_t, size_t, hipStream_t);

template hipError_t gen_post_twiddles<float>(C2<float> *, C2<float> *, int, size_t, hipStream_t);
template hipError_t gen_post_twiddles<double>(C2<double> *, C2<double> *, int, size_t,
                                              hipStream_t);

// -------------------------------------------------------- synthetic inputs --

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ __forceinline__ uint64_t fill_key(uint64_t seed) {
    uint64_t z = seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__device__ __forceinline__ void fill_range(T *dst, uint64_t count, int dist, uint64_t key,
                                           uint64_t first, uint64_t stride) {
    for (uint64_t i = first; i < count; i += stride) {
        if (dist == 0) {
            const uint64_t w = mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
            const double u = (double)(w >> 40) * (1.0 / 16777216.0);
            dst[i] = (T)(2.0 * u - 1.0);
        } else {
            // Box-Muller on words (2p, 2p+1) of pair p = i/2; even i takes cos, odd i sin.
            const uint64_t p = i >> 1;
            const uint64_t w1 = mix64(key + (2 * p + 1) * 0x9E3779B97F4A7C15ull);
            const uint64_t w2 = mix64(key + (2 * p + 2) * 0x9E3779B97F4A7C15ull);
            const double u1 = ((double)(w1 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
            const double u2 = (double)(w2 >> 11) * (1.0 / 9007199254740992.0);
            const double r = sqrt(-2.0 * log(u1));
            double s, c;
            sincospi(2.0 * u2, &s, &c);
            dst[i] = (T)((i & 1) ? r * s : r * c);
        }
    }
}

template <typename T>
__global__ void k_fill(T *dst, uint64_t count, int dist, uint64_t key) {
    fill_range(dst, count, dist, key, blockIdx.x * (uint64_t)blockDim.x + threadIdx.x,
               (uint64_t)gridDim.x * blockDim.x);
}

// Uniform [-1, 1) value of word t = key + (i + 1) * golden, as fill_range
// computes it, in 32-bit arithmetic: only the top 24 bits of the word are
// used, and the final xor-shift of the finaliser leaves them unchanged
// (z >> 31 is below 2^33), so they are the top bits of the second product,
// whose high half needs no low product.  m * 2^-24 * 2 - 1 = (m - 2^23) * 2^-23
// is exact in FP32 and FP64.
template <typename T>
__device__ __forceinline__ T uniform_pm1(uint64_t t) {
    t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
    t ^= t >> 27;
    const uint32_t lo = (uint32_t)t, hi = (uint32_t)(t >> 32);
    const uint32_t top = __umulhi(lo, 0x133111EBu) + lo * 0x94D049BBu + hi * 0x133111EBu;
    const int m = (int)(top >> 8) - (1 << 23);
    return (T)m * (T)(1.0 / 8388608.0);
}

// Uniform fill with 16-byte stores: each thread writes V consecutive values
// per step; the unaligned head (dst is T-aligned) and the tail are written
// by the first threads one value at a time.
template <typename T>
__device__ __forceinline__ void fill_uniform_range(T *dst, uint64_t count, uint64_t key,
                                                   uint64_t tid, uint64_t stride) {
    constexpr int V = 16 / sizeof(T);
    constexpr uint64_t G = 0x9E3779B97F4A7C15ull;
    const uint64_t mis = ((uintptr_t)dst & 15) / sizeof(T);
    const uint64_t want = mis ? (uint64_t)V - mis : 0, head = want < count ? want : count;
    const uint64_t nv = (count - head) / V, tail = head + nv * V;
    if (tid < head) dst[tid] = uniform_pm1<T>(key + (tid + 1) * G);
    if (tid < count - tail) dst[tail + tid] = uniform_pm1<T>(key + (tail + tid + 1) * G);
    using Vec = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
    Vec *body = reinterpret_cast<Vec *>(dst + head);
    uint64_t t = key + (head + tid * V + 1) * G;
    const uint64_t step = stride * V * G;
    for (uint64_t q = tid; q < nv; q += stride, t += step) {
        T v[V];
#pragma unroll
        for (int k = 0; k < V; ++k) v[k] = uniform_pm1<T>(t + (uint64_t)k * G);
        Vec o;
        __builtin_memcpy(&o, v, sizeof(o));
        body[q] = o;
    }
}

template <typename T>
__global__ void k_fill_uniform(T *dst, uint64_t count, uint64_t key) {
    fill_uniform_range(dst, count, key, blockIdx.x * (uint64_t)blockDim.x + threadIdx.x,
                       (uint64_t)gridDim.x * blockDim.x);
}

// One buffer per blockIdx.y.
__global__ void k_fill_many(const FillEntry *entries) {
    const FillEntry e = entries[blockIdx.y];
    const uint64_t first = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t key = fill_key(e.seed);
    if (e.dist == 0) {
        if (e.f64)
            fill_uniform_range(static_cast<double *>(e.dst), e.count, key, first, stride);
        else
            fill_uniform_range(static_cast<float *>(e.dst), e.count, key, first, stride);
    } else if (e.f64) {
        fill_range(static_cast<double *>(e.dst), e.count, e.dist, key, first, stride);
    } else {
        fill_range(static_cast<float *>(e.dst), e.count, e.dist, key, first, stride);
    }
}

// workgroups of 256 threads for `count` values, 16 bytes per thread and step
static unsigned fill_blocks(uint64_t count, int f64) {
    const uint64_t per_block = 256ull * (f64 ? 2 : 4);
    const uint64_t b = (count + per_block - 1) / per_block;
    return (unsigned)(b < 1 ? 1 : b > 16384 ? 16384 : b);
}

hipError_t fill_synthetic_many(const FillEntry *entries, int n, uint64_t max_count, int f64,
                               hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_many, dim3(fill_blocks(max_count, f64), (unsigned)n), dim3(256), 0,
                       stream, entries);
    return hipGetLastError();
}

hipError_t fill_synthetic(void *dst, size_t count, int f64, int dist, uint64_t seed,
                          hipStream_t stream) {
    if (count == 0) return hipSuccess;
    const uint64_t key = fill_key(seed);
    if (dist == 0 && (uintptr_t)dst % (f64 ? 8 : 4) == 0) {
        const unsigned blocks = fill_blocks(count, f64);
        if (f64)
            hipLaunchKernelGGL((k_fill_uniform<double>), dim3(blocks), dim3(256), 0, stream,
                               (double *)dst, (uint64_t)count, key);
        else
            hipLaunchKernelGGL((k_fill_uniform<float>), dim3(blocks), dim3(256), 0, stream,
                               (float *)dst, (uint64_t)count, key);
        return hipGetLastError();
    }
    size_t blocks = (count + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (f64)
        hipLaunchKernelGGL((k_fill<double>), dim3((unsigned)blocks), dim3(256), 0, stream,
                           (double *)dst, (uint64_t)count, dist, key);
    else
        hipLaunchKernelGGL((k_fill<float>), dim3((unsigned)blocks), dim3(256), 0, stream,
                           (float *)dst, (uint64_t)count, dist, key);
    return hipGetLastError();
}

}  // namespace sfft
