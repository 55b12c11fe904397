// fixed.h -- compile-time planned, persistent, software-pipelined transforms.
//
// For the lengths that carry most traffic the radix sequence, spans, thread
// layout and LDS offsets are template constants: index arithmetic folds into
// immediates and each kernel holds only its own stages.  The kernels compute
// exactly what the runtime-planned pass kernel computes for the same radices
// (same butterflies, same twiddle table, explicit FMAs), so a plan gives the
// same bits on either path.
//
// Persistent + pipelined: a block walks slot groups blockIdx.x,
// blockIdx.x + gridDim.x, ...; the first stage of a group takes its elements
// from registers loaded during the previous group, and right after that the
// next group's loads are issued, so HBM reads overlap the remaining stages.
#pragma once

#include "stockham.h"

namespace sfft {

constexpr bool ct_pow2(int x) { return x > 0 && (x & (x - 1)) == 0; }
constexpr int ct_pad(int e) { return e + (e >> kLdsPadShift); }

template <typename T>
struct FixedArgs {
    const C2<T> *in;
    C2<T> *out;
    const C2<T> *tw;       // stage twiddles (dense layout of the runtime planner)
    const C2<T> *real_tw;  // twiddlesMinusI, N/4 + 1 entries of the real length
    const C2<T> *rot;      // half-shift rotations
    int32_t total_slots;
    uint32_t flip_in;      // sign-bit mask xor-ed into Im at the HBM load (inverse)
    uint32_t flip_out;     // ... and at the HBM store
    int32_t half_shift;
};

// Sign flip of the imaginary part by a uniform xor mask (0 or the sign bit).
__device__ __forceinline__ void flip_im(C2<float> &x, uint32_t m) {
    x.y = __uint_as_float(__float_as_uint(x.y) ^ m);
}
__device__ __forceinline__ void flip_im(C2<double> &x, uint32_t m) {
    unsigned long long u = __double_as_longlong(x.y);
    u ^= (unsigned long long)m << 32;
    x.y = __longlong_as_double(u);
}

template <typename T_, int N, int TPT, int TPW, int MODE, int... Rs>
struct Fixed {
    using T = T_;
    static constexpr int NST = sizeof...(Rs);
    static constexpr int RAD[NST] = {Rs...};
    static constexpr int THREADS = TPT * TPW;
    static constexpr int STRIDE = N + (N >> kLdsPadShift) + 1;  // lds_padded(N)
    static constexpr uint32_t EB = sizeof(C2<T>);
    static constexpr size_t LDS_BYTES =
        (NST > 1 || MODE != kModeC2C) ? (size_t)STRIDE * EB * TPW : 0;

    static constexpr int ns_of(int s) {
        int p = 1;
        for (int i = 0; i < s; ++i) p *= RAD[i];
        return p;
    }
    static constexpr int tw_of(int s) {
        int o = 0;
        for (int i = 0; i < s; ++i)
            if (ns_of(i) > 1) o += ns_of(i) * (RAD[i] - 1);
        return o;
    }
    static constexpr int bfly_of(int s) { return (N / RAD[s] + TPT - 1) / TPT; }
    static constexpr int PF = bfly_of(0) * RAD[0];
    static_assert(ns_of(NST) == N, "radices must multiply to N");

    // stage-0 element loads of slot g into pf (ROW mapping, contiguous batch)
    static __device__ __forceinline__ void load0(const FixedArgs<T> &a, C2<T> (&pf)[PF], int g,
                                                 int lid) {
        constexpr int R = RAD[0], NB = N / R, B = bfly_of(0);
        if (g >= a.total_slots) return;
        const char *src = reinterpret_cast<const char *>(a.in);
        const uint32_t lane = ((uint32_t)g * (uint32_t)N + (uint32_t)lid) * EB;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if ((NB % TPT) == 0 || lid + b * TPT < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    pf[b * R + r] = *reinterpret_cast<const C2<T> *>(
                        src + (lane + (uint32_t)((b * TPT + r * NB) * EB)));
            }
        }
    }

    template <int S>
    static __device__ __forceinline__ void stage(const FixedArgs<T> &a, char *lds_c, int g,
                                                 bool active, int slot, int lid, C2<T> (&pf)[PF],
                                                 int g_next) {
        constexpr int R = RAD[S], NS = ns_of(S), NB = N / R, B = bfly_of(S);
        constexpr bool FIRST = (S == 0), LAST = (S == NST - 1);
        constexpr bool SRC_PF = FIRST && MODE != kModeC2R;
        constexpr bool DST_G = LAST && MODE != kModeR2C;
        constexpr bool SEP_R = ct_pow2(NB) || (NB % 16) == 0;
        constexpr bool SEP_W = (NS % 16) == 0 || (ct_pow2(NS) && ct_pow2(R));
        constexpr int TWO = tw_of(S);
        const uint32_t lsb = (uint32_t)(slot * STRIDE) * EB;

        auto jvalid = [&](int b) { return (NB % TPT) == 0 || lid + b * TPT < NB; };

        C2<T> v[B][R];
        // twiddles (issued before the LDS phase; 16 or fewer per thread)
        C2<T> w[B][R];
        if constexpr (NS > 1) {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    const int k = j % NS;
                    const C2<T> *wp = a.tw + (TWO + k * (R - 1));
#pragma unroll
                    for (int r = 1; r < R; ++r) w[b][r] = wp[r - 1];
                }
            }
        }
        if constexpr (SRC_PF) {
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    v[b][r] = pf[b * R + r];
                    flip_im(v[b][r], a.flip_in);
                }
            if (MODE == kModeR2C && a.half_shift) {
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = lid + b * TPT;
                    if (active && jvalid(b)) {
#pragma unroll
                        for (int r = 0; r < R; ++r) v[b][r] = cmul(v[b][r], a.rot[j + r * NB]);
                    }
                }
            }
            if (g_next >= 0) load0(a, pf, g_next, lid);
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    if constexpr (SEP_R) {
                        const uint32_t la = lsb + (uint32_t)ct_pad(j) * EB;
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            v[b][r] = *reinterpret_cast<const C2<T> *>(
                                lds_c + (la + (uint32_t)(ct_pad(r * NB) * EB)));
                    } else {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            v[b][r] = *reinterpret_cast<const C2<T> *>(
                                lds_c + (lsb + (uint32_t)ct_pad(j + r * NB) * EB));
                    }
                }
            }
            __syncthreads();  // all reads done before any slot is overwritten
        }
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if (active && jvalid(b)) {
                if constexpr (NS > 1) {
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[b][r]);
                }
                dft<R, T>(v[b]);
            }
        }
        if constexpr (DST_G) {
            char *dst = reinterpret_cast<char *>(a.out);
            const uint32_t sbase = (uint32_t)g * (uint32_t)N;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    const int o = (j / NS) * NS * R + j % NS;
                    const uint32_t lo = (sbase + (uint32_t)o) * EB;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        C2<T> y = v[b][t];
                        flip_im(y, a.flip_out);
                        if (MODE == kModeC2R && a.half_shift) y = cmulc(y, a.rot[o + t * NS]);
                        *reinterpret_cast<C2<T> *>(dst + (lo + (uint32_t)(t * NS * EB))) = y;
                    }
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    const int o = (j / NS) * NS * R + j % NS;
                    if constexpr (SEP_W) {
                        const uint32_t la = lsb + (uint32_t)ct_pad(o) * EB;
#pragma unroll
                        for (int t = 0; t < R; ++t)
                            *reinterpret_cast<C2<T> *>(
                                lds_c + (la + (uint32_t)(ct_pad(t * NS) * EB))) = v[b][t];
                    } else {
#pragma unroll
                        for (int t = 0; t < R; ++t)
                            *reinterpret_cast<C2<T> *>(
                                lds_c + (lsb + (uint32_t)ct_pad(o + t * NS) * EB)) = v[b][t];
                    }
                }
            }
            __syncthreads();
        }
    }

    template <int S>
    static __device__ __forceinline__ void stages_from(const FixedArgs<T> &a, char *lds_c, int g,
                                                       bool active, int slot, int lid,
                                                       C2<T> (&pf)[PF], int g_next) {
        if constexpr (S < NST) {
            stage<S>(a, lds_c, g, active, slot, lid, pf, g_next);
            stages_from<S + 1>(a, lds_c, g, active, slot, lid, pf, g_next);
        }
    }

    // Real inverse pre-process (pairs i, h - i) from HBM into LDS, storing
    // conj(.) so the forward stages compute the inverse (signalsmith-fft.h:478-492).
    static __device__ __forceinline__ void real_pre(const FixedArgs<T> &a, char *lds_c, int g,
                                                    bool active, int slot, int lid) {
        constexpr int H = N;
        C2<T> *ls = reinterpret_cast<C2<T> *>(lds_c) + slot * STRIDE;
        const bool mod = a.half_shift != 0;
        if (active) {
            const C2<T> *x = a.in + (size_t)g * H;
            if (!mod && lid == 0) {
                const C2<T> v = x[0];
                ls[0] = {v.x + v.y, -(v.x - v.y)};
            }
            for (int i = (mod ? 0 : 1) + lid; i <= H / 2; i += TPT) {
                const int ci = mod ? (H - 1 - i) : (H - i);
                const C2<T> v = x[i], v2 = x[ci];
                const C2<T> odd = {v.x + v2.x, v.y - v2.y};
                const C2<T> erot = {v.x - v2.x, v.y + v2.y};
                const C2<T> ev = cmulc(erot, a.real_tw[i]);
                const C2<T> p = cadd(odd, ev);
                const C2<T> m = csub(odd, ev);
                ls[ct_pad(i)] = cconj(p);
                ls[ct_pad(ci)] = m;
            }
        }
        __syncthreads();
    }

    // Real forward post-process from the spectrum in LDS to HBM
    // (signalsmith-fft.h:459-472).
    static __device__ __forceinline__ void real_post(const FixedArgs<T> &a, char *lds_c, int g,
                                                     bool active, int slot, int lid) {
        constexpr int H = N;
        const C2<T> *ls = reinterpret_cast<const C2<T> *>(lds_c) + slot * STRIDE;
        const bool mod = a.half_shift != 0;
        if (active) {
            C2<T> *y = a.out + (size_t)g * H;
            if (!mod && lid == 0) {
                const C2<T> z = ls[0];
                y[0] = {z.x + z.y, z.x - z.y};
            }
            const T half = (T)0.5;
            for (int i = (mod ? 0 : 1) + lid; i <= H / 2; i += TPT) {
                const int ci = mod ? (H - 1 - i) : (H - i);
                const C2<T> zi = ls[ct_pad(i)], zc = ls[ct_pad(ci)];
                const C2<T> odd = {(zi.x + zc.x) * half, (zi.y - zc.y) * half};
                const C2<T> ev = {(zi.x - zc.x) * half, (zi.y + zc.y) * half};
                const C2<T> e = cmul(ev, a.real_tw[i]);
                y[i] = cadd(odd, e);
                y[ci] = cconj(csub(odd, e));
            }
        }
        __syncthreads();
    }

    static __device__ __forceinline__ void run(const FixedArgs<T> &a, char *lds_c) {
        const int tid = (int)threadIdx.x;
        const int slot = tid / TPT, lid = tid % TPT;
        const int ngroups = (a.total_slots + TPW - 1) / TPW;
        C2<T> pf[PF];
        int grp = (int)blockIdx.x;
        if (MODE != kModeC2R && grp < ngroups) load0(a, pf, grp * TPW + slot, lid);
        for (; grp < ngroups; grp += (int)gridDim.x) {
            const int g = grp * TPW + slot;
            const bool active = g < a.total_slots;
            const int gn = grp + (int)gridDim.x;
            const int g_next = gn < ngroups ? gn * TPW + slot : -1;
            if constexpr (MODE == kModeC2R) real_pre(a, lds_c, g, active, slot, lid);
            stages_from<0>(a, lds_c, g, active, slot, lid, pf, g_next);
            if constexpr (MODE == kModeR2C) real_post(a, lds_c, g, active, slot, lid);
        }
    }
};

template <class F>
__global__ void __launch_bounds__(F::THREADS) k_fixed(FixedArgs<typename F::T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    F::run(a, reinterpret_cast<char *>(smem_raw));
}

// Registry entry of a compiled fixed kernel.
struct FixedKernel {
    int n;          // complex length of the core transform
    int prec;       // 0 F32, 1 F64
    int real;       // 0: C2C kernel; 1: real pair (R2C forward + C2R inverse)
    int tpt, tpw;
    int nst;
    int radices[kMaxStages];
    size_t lds_bytes;
    // launch(args, mode, grid, stream): mode kModeC2C / kModeR2C / kModeC2R
    hipError_t (*launch)(const void *args, int mode, unsigned grid, hipStream_t stream);
    int (*blocks_per_cu)(int mode);
};

void register_fixed(const FixedKernel &k);
const FixedKernel *find_fixed(int n, int prec, int real);

template <class F>
hipError_t fixed_launch_one(const void *args, unsigned grid, hipStream_t stream) {
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void *)k_fixed<F>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize,
                                           160 * 1024 - kLdsReserve);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return e;
        }
        attr = true;
    }
    hipLaunchKernelGGL((k_fixed<F>), dim3(grid), dim3(F::THREADS), F::LDS_BYTES, stream,
                       *reinterpret_cast<const FixedArgs<typename F::T> *>(args));
    return hipGetLastError();
}

template <class F>
int fixed_occupancy() {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)k_fixed<F>,
                                                     F::THREADS, F::LDS_BYTES) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return blocks > 0 ? blocks : 1;
}

template <int... Rs>
inline void fixed_fill_radices(FixedKernel &k) {
    k.nst = (int)sizeof...(Rs);
    const int rs[] = {Rs...};
    for (int i = 0; i < k.nst; ++i) k.radices[i] = rs[i];
}

// Registers the complex kernel of one compile-time plan (namespace scope
// object in a .hip file).
template <typename T, int N, int TPT, int TPW, int... Rs>
struct FixedC2C {
    using F = Fixed<T, N, TPT, TPW, kModeC2C, Rs...>;
    FixedC2C() {
        FixedKernel k{};
        k.n = N;
        k.prec = sizeof(T) == 8 ? 1 : 0;
        k.real = 0;
        k.tpt = TPT;
        k.tpw = TPW;
        fixed_fill_radices<Rs...>(k);
        k.lds_bytes = F::LDS_BYTES;
        k.launch = [](const void *args, int, unsigned grid, hipStream_t s) {
            return fixed_launch_one<F>(args, grid, s);
        };
        k.blocks_per_cu = [](int) { return fixed_occupancy<F>(); };
        register_fixed(k);
    }
};

// Registers the real forward / inverse pair of one compile-time plan of the
// half-length complex core N (real length 2N).
template <typename T, int N, int TPT, int TPW, int... Rs>
struct FixedReal {
    using FR = Fixed<T, N, TPT, TPW, kModeR2C, Rs...>;
    using FI = Fixed<T, N, TPT, TPW, kModeC2R, Rs...>;
    FixedReal() {
        FixedKernel k{};
        k.n = N;
        k.prec = sizeof(T) == 8 ? 1 : 0;
        k.real = 1;
        k.tpt = TPT;
        k.tpw = TPW;
        fixed_fill_radices<Rs...>(k);
        k.lds_bytes = FR::LDS_BYTES;
        k.launch = [](const void *args, int mode, unsigned grid, hipStream_t s) {
            return mode == kModeC2R ? fixed_launch_one<FI>(args, grid, s)
                                    : fixed_launch_one<FR>(args, grid, s);
        };
        k.blocks_per_cu = [](int mode) {
            return mode == kModeC2R ? fixed_occupancy<FI>() : fixed_occupancy<FR>();
        };
        register_fixed(k);
    }
};

}  // namespace sfft
