// storage.h -- HBM storage formats and the radix-16 MFMA contraction.
//
// The arithmetic type of a transform is T (float or double).  Its HBM data
// may be stored in T itself or, for T = float, in 16-bit formats (IEEE half
// or bfloat16, interleaved (re, im), 4 bytes per complex value): loads widen
// to float, stores round to nearest even.  LDS and registers always hold T.
//
// With 16-bit storage the radix-16 butterflies run on the matrix cores: a
// 16-point DFT of 32 columns is the real 32x32 product
//     [Yr; Yi] = [[C, S], [-S, C]] [Xr; Xi],  C = cos(2 pi m c / 16),
//                                              S = sin(2 pi m c / 16)
// done as four v_mfma_f32_32x32x8 steps (K = 8 real rows each) accumulated
// in float.  Operands are split into two 16-bit parts (x = hi + lo, both for
// the data and for the DFT matrix) and the three products hi*hi, hi*lo,
// lo*hi are accumulated, which keeps the contraction near float accuracy
// (the lo*lo term is below 2^-21 of the result).  Each column is scaled by a
// power of two before the split so that half's range is never exceeded.
//
// Lane layout of v_mfma_f32_32x32x8 (checked on gfx950): lane l holds
// A[l % 32][4 (l / 32) + i] and B[4 (l / 32) + i][l % 32] for i < 4, and the
// result D[8 (a / 4) + 4 (l / 32) + (a % 4)][l % 32] in accumulator a < 16.
// With real rows 0..15 = re, 16..31 = im, lane l of column l % 32 therefore
// holds points {4h..4h+3, 8+4h..8+4h+3} (h = l / 32) of its butterfly, both
// as input and as output.
#pragma once

#include <stdint.h>

#include "dft.h"

namespace sfft {

struct StoreNative {};
struct StoreF16 {};
struct StoreBF16 {};

// float -> bfloat16 bits, round to nearest even (NaN stays NaN).
__device__ __forceinline__ uint32_t f32_to_bf16(float x) {
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u >> 16) | 0x40u;
    u += 0x7fffu + ((u >> 16) & 1u);
    return u >> 16;
}
__device__ __forceinline__ uint32_t f32_to_f16(float x) {
    return (uint32_t)__builtin_bit_cast(uint16_t, (_Float16)x);
}
__device__ __forceinline__ float f16_to_f32(uint32_t b) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)b);
}

template <typename T, typename S>
struct Storage;

template <typename T>
struct Storage<T, StoreNative> {
    static constexpr bool native = true;
    __device__ static C2<T> load(const void *p, int64_t i) {
        return static_cast<const C2<T> *>(p)[i];
    }
    __device__ static void store(void *p, int64_t i, C2<T> v) { static_cast<C2<T> *>(p)[i] = v; }
};

template <>
struct Storage<float, StoreF16> {
    static constexpr bool native = false;
    __device__ static C2<float> load(const void *p, int64_t i) {
        const uint32_t w = static_cast<const uint32_t *>(p)[i];
        return {f16_to_f32(w & 0xffffu), f16_to_f32(w >> 16)};
    }
    __device__ static void store(void *p, int64_t i, C2<float> v) {
        static_cast<uint32_t *>(p)[i] = f32_to_f16(v.x) | (f32_to_f16(v.y) << 16);
    }
};

template <>
struct Storage<float, StoreBF16> {
    static constexpr bool native = false;
    __device__ static C2<float> load(const void *p, int64_t i) {
        const uint32_t w = static_cast<const uint32_t *>(p)[i];
        return {__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
    }
    __device__ static void store(void *p, int64_t i, C2<float> v) {
        static_cast<uint32_t *>(p)[i] = f32_to_bf16(v.x) | (f32_to_bf16(v.y) << 16);
    }
};

typedef float f32x16 __attribute__((ext_vector_type(16)));

template <typename S>
struct Mfma16;

template <>
struct Mfma16<StoreF16> {
    typedef _Float16 v4 __attribute__((ext_vector_type(4)));
    typedef _Float16 elem;
    __device__ static void split(float x, elem &hi, elem &lo) {
        hi = (_Float16)x;
        lo = (_Float16)(x - (float)hi);
    }
    __device__ static f32x16 mma(v4 a, v4 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x8f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma16<StoreBF16> {
    typedef short v4 __attribute__((ext_vector_type(4)));
    typedef short elem;
    __device__ static void split(float x, elem &hi, elem &lo) {
        const uint32_t h = f32_to_bf16(x);
        hi = (short)h;
        lo = (short)f32_to_bf16(x - __uint_as_float(h << 16));
    }
    __device__ static f32x16 mma(v4 a, v4 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, c, 0, 0, 0);
    }
};

// cos(2 pi t / 16), t = 0..15.
static __constant__ float kCos16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};

// Point index r of local value i (0..7) held by lane half h (0 or 1).
__device__ __forceinline__ int mfma_point(int i, int h) { return (i < 4) ? 4 * h + i : 8 + 4 * h + (i - 4); }

// The DFT-16 matrix operands of this lane, split hi + lo: for K step ks the
// lane holds G[m][8 ks + 4 h + i], m = lane % 32, i < 4.
template <typename S>
struct Dft16Operands {
    typename Mfma16<S>::v4 hi[4], lo[4];
    __device__ Dft16Operands(int lane) {
        const int m = lane & 31, h = lane >> 5;
        const int mm = m & 15;
        const bool out_im = m >= 16;
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int c = 8 * ks + 4 * h + i;
                const int cc = c & 15;
                const bool in_im = c >= 16;
                const int t = (mm * cc) & 15;
                const float cs = kCos16[t], sn = kCos16[(t + 12) & 15];  // sin(x) = cos(x - pi/2)
                // [[C, S], [-S, C]]
                const float g = (out_im == in_im) ? cs : (in_im ? sn : -sn);
                typename Mfma16<S>::elem a, b;
                Mfma16<S>::split(g, a, b);
                hi[ks][i] = a;
                lo[ks][i] = b;
            }
        }
    }
};

// Forward 16-point DFTs of the 32 columns held by a wave (8 points per lane,
// local order of mfma_point), in place.
template <typename S>
__device__ __forceinline__ void dft16_mfma(C2<float> (&v)[8], const Dft16Operands<S> &g) {
    using M = Mfma16<S>;
    float mx = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) mx = fmaxf(mx, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
    mx = fmaxf(mx, __shfl_xor(mx, 32));
    int ex = 0;
    (void)frexpf(mx, &ex);  // mx < 2^ex: the scaled column lies in (-1, 1)
    typename M::v4 bh[4], bl[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        typename M::elem h, l;
        M::split(ldexpf(v[i].x, -ex), h, l);
        bh[0][i] = h;
        bl[0][i] = l;
        M::split(ldexpf(v[4 + i].x, -ex), h, l);
        bh[1][i] = h;
        bl[1][i] = l;
        M::split(ldexpf(v[i].y, -ex), h, l);
        bh[2][i] = h;
        bl[2][i] = l;
        M::split(ldexpf(v[4 + i].y, -ex), h, l);
        bh[3][i] = h;
        bl[3][i] = l;
    }
    f32x16 acc = {};
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
        acc = M::mma(g.hi[ks], bh[ks], acc);
        acc = M::mma(g.hi[ks], bl[ks], acc);
        acc = M::mma(g.lo[ks], bh[ks], acc);
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) v[i] = {ldexpf(acc[i], ex), ldexpf(acc[8 + i], ex)};
}

}  // namespace sfft
