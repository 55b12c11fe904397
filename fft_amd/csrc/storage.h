// storage.h -- HBM storage formats and the radix-16 MFMA contraction.
//
// The arithmetic type of a transform is T (float or double).  Its HBM data
// may be stored in T itself or, for T = float, in 16-bit formats (IEEE half
// or bfloat16, interleaved (re, im), 4 bytes per complex value), or, for
// T = double, in FP32 (8 bytes per complex value): loads widen, stores round
// to nearest even.  LDS and registers always hold T.
//
// With 16-bit storage the radix-16 butterflies run on the matrix cores: a
// 16-point DFT of 32 columns is the real 32x32 product
//     [Yr; Yi] = [[C, S], [-S, C]] [Xr; Xi],  C = cos(2 pi m c / 16),
//                                              S = sin(2 pi m c / 16)
// done as two gfx950 v_mfma_f32_32x32x16 steps (K = 16 real rows each)
// accumulated in float.  Operands are split into two 16-bit parts (x = hi +
// lo, both for the data and for the DFT matrix) and the three products
// hi*hi, hi*lo, lo*hi are accumulated, which keeps the contraction near float
// accuracy (the lo*lo term is below 2^-21 of the result).  With IEEE half
// each column is scaled by a power of two before the split so that half's
// range is never exceeded; bfloat16 has float's exponent range and needs none.
//
// Lane layout of v_mfma_f32_32x32x16: lane l holds A[l % 32][8 (l / 32) + i]
// and B[8 (l / 32) + i][l % 32] for i < 8, and the result
// D[8 (a / 4) + 4 (l / 32) + (a % 4)][l % 32] in accumulator a < 16.  With
// real rows 0..15 = re, 16..31 = im, lane l of column l % 32 (h = l / 32)
// therefore takes input points 8h..8h+7 of its butterfly (mfma_in_point) and
// holds output points {4h..4h+3, 8+4h..8+4h+3} (mfma_point).
#pragma once

#include <stdint.h>

#include "dft.h"

namespace sfft {

struct StoreNative {};
struct StoreF16 {};
struct StoreBF16 {};
struct StoreF32 {};  // FP32 complex values in HBM under FP64 arithmetic

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
    static constexpr int kBytes = (int)sizeof(C2<T>);  // bytes per complex value in HBM
    __device__ static C2<T> load(const void *p, int64_t i) {
        return static_cast<const C2<T> *>(p)[i];
    }
    // a load marked non-temporal (data read once)
    __device__ static C2<T> load_nt(const void *p, int64_t i) {
        using V2 = T __attribute__((ext_vector_type(2)));
        const V2 v = __builtin_nontemporal_load(static_cast<const V2 *>(p) + i);
        return {v.x, v.y};
    }
    __device__ static void store(void *p, int64_t i, C2<T> v) { static_cast<C2<T> *>(p)[i] = v; }
};

template <>
struct Storage<float, StoreF16> {
    static constexpr bool native = false;
    static constexpr int kBytes = 4;
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
    static constexpr int kBytes = 4;
    __device__ static C2<float> load(const void *p, int64_t i) {
        const uint32_t w = static_cast<const uint32_t *>(p)[i];
        return {__uint_as_float(w << 16), __uint_as_float(w & 0xffff0000u)};
    }
    __device__ static void store(void *p, int64_t i, C2<float> v) {
        typedef __bf16 b2 __attribute__((ext_vector_type(2)));
        const b2 w = {(__bf16)v.x, (__bf16)v.y};  // one packed round-to-nearest-even convert
        static_cast<uint32_t *>(p)[i] = __builtin_bit_cast(uint32_t, w);
    }
};

// FP32 storage of an FP64 transform: loads widen exactly, stores round to
// nearest even once.
template <>
struct Storage<double, StoreF32> {
    static constexpr bool native = false;
    static constexpr int kBytes = 8;
    __device__ static C2<double> load(const void *p, int64_t i) {
        const C2<float> v = static_cast<const C2<float> *>(p)[i];
        return {(double)v.x, (double)v.y};
    }
    __device__ static void store(void *p, int64_t i, C2<double> v) {
        static_cast<C2<float> *>(p)[i] = {(float)v.x, (float)v.y};
    }
};

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

// split2(a, b): the 16-bit hi parts of a and b packed in one dword (a in the
// low half) and the 16-bit parts of the residuals a - hi, b - hi likewise.
// The residual of a 16-bit rounding is exact in float, so x = hi + lo up to
// the rounding of lo (2^-11 / 2^-8 of lo, about 2^-22 / 2^-16 of x).
template <typename S>
struct Mfma16;

template <>
struct Mfma16<StoreF16> {
    typedef _Float16 v8 __attribute__((ext_vector_type(8)));
    static constexpr bool kScale = true;  // half's range: scale each column
    __device__ static void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
        const auto h = __builtin_amdgcn_cvt_pkrtz(a, b);
        const auto l = __builtin_amdgcn_cvt_pkrtz(a - (float)h[0], b - (float)h[1]);
        hi = __builtin_bit_cast(uint32_t, h);
        lo = __builtin_bit_cast(uint32_t, l);
    }
    __device__ static f32x16 mma(v8 a, v8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
    }
};

template <>
struct Mfma16<StoreBF16> {
    typedef __bf16 v8 __attribute__((ext_vector_type(8)));
    typedef __bf16 b2 __attribute__((ext_vector_type(2)));
    static constexpr bool kScale = false;  // float's exponent range
    __device__ static void split2(float a, float b, uint32_t &hi, uint32_t &lo) {
        const b2 h = {(__bf16)a, (__bf16)b};
        const uint32_t hw = __builtin_bit_cast(uint32_t, h);
        const b2 l = {(__bf16)(a - __uint_as_float(hw << 16)),
                      (__bf16)(b - __uint_as_float(hw & 0xffff0000u))};
        hi = hw;
        lo = __builtin_bit_cast(uint32_t, l);
    }
    __device__ static f32x16 mma(v8 a, v8 b, f32x16 c) {
        return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
    }
};

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <typename S>
__device__ __forceinline__ typename Mfma16<S>::v8 as_v8(uint32_t w0, uint32_t w1, uint32_t w2,
                                                       uint32_t w3) {
    return __builtin_bit_cast(typename Mfma16<S>::v8, (u32x4){w0, w1, w2, w3});
}

// cos(2 pi t / 16), t = 0..15.
static __constant__ float kCos16[16] = {
    1.0f, 0.92387953251128674f, 0.70710678118654752f, 0.38268343236508977f,
    0.0f, -0.38268343236508977f, -0.70710678118654752f, -0.92387953251128674f,
    -1.0f, -0.92387953251128674f, -0.70710678118654752f, -0.38268343236508977f,
    0.0f, 0.38268343236508977f, 0.70710678118654752f, 0.92387953251128674f};

// Output point index of local value i (0..7) held by lane half h (0 or 1).
__device__ __forceinline__ int mfma_point(int i, int h) { return (i < 4) ? 4 * h + i : 8 + 4 * h + (i - 4); }
// Input point index of local value i held by lane half h.
__device__ __forceinline__ int mfma_in_point(int i, int h) { return 8 * h + i; }

// The DFT-16 matrix operands of this lane, split hi + lo: for K step ks the
// lane holds G[m][16 ks + 8 h + i], m = lane % 32, i < 8.
template <typename S>
struct Dft16Operands {
    typename Mfma16<S>::v8 hi[2], lo[2];
    __device__ static float entry(int m, int c) {
        const int mm = m & 15, cc = c & 15;
        const bool out_im = m >= 16, in_im = c >= 16;
        const int t = (mm * cc) & 15;
        const float cs = kCos16[t], sn = kCos16[(t + 12) & 15];  // sin(x) = cos(x - pi/2)
        return (out_im == in_im) ? cs : (in_im ? sn : -sn);     // [[C, S], [-S, C]]
    }
    __device__ explicit Dft16Operands(int lane) {
        const int m = lane & 31, h = lane >> 5;
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const int c0 = 16 * ks + 8 * h;
            uint32_t hw[4], lw[4];
#pragma unroll
            for (int p = 0; p < 4; ++p)
                Mfma16<S>::split2(entry(m, c0 + 2 * p), entry(m, c0 + 2 * p + 1), hw[p], lw[p]);
            hi[ks] = as_v8<S>(hw[0], hw[1], hw[2], hw[3]);
            lo[ks] = as_v8<S>(lw[0], lw[1], lw[2], lw[3]);
        }
    }
};

// Forward 16-point DFTs of G groups of 32 columns held by a wave (8 points
// per lane and group: input points in the order of mfma_in_point, output
// points in the order of mfma_point), in place.  K steps: ks 0 = re of the
// lane's input points, 1 = their im (rows 16 ks + 8 h + i).  The MFMAs of
// the groups are interleaved, and the hi*hi products and the two correction
// products accumulate separately, so 2 G independent accumulation chains are
// in flight.
template <typename S, int G>
__device__ __forceinline__ void dft16_mfma(C2<float> (&v)[G][8], const Dft16Operands<S> &g) {
    using M = Mfma16<S>;
    int ex[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        ex[q] = 0;
        if constexpr (M::kScale) {
            float mx = 0.f;
#pragma unroll
            for (int i = 0; i < 8; ++i) mx = fmaxf(mx, fmaxf(fabsf(v[q][i].x), fabsf(v[q][i].y)));
            mx = fmaxf(mx, __shfl_xor(mx, 32));
            (void)frexpf(mx, &ex[q]);  // mx < 2^ex: the scaled column lies in (-1, 1)
#pragma unroll
            for (int i = 0; i < 8; ++i)
                v[q][i] = {ldexpf(v[q][i].x, -ex[q]), ldexpf(v[q][i].y, -ex[q])};
        }
    }
    typename M::v8 bh[G][2], bl[G][2];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        uint32_t h[2][4], l[2][4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
            M::split2(v[q][2 * p].x, v[q][2 * p + 1].x, h[0][p], l[0][p]);
            M::split2(v[q][2 * p].y, v[q][2 * p + 1].y, h[1][p], l[1][p]);
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            bh[q][ks] = as_v8<S>(h[ks][0], h[ks][1], h[ks][2], h[ks][3]);
            bl[q][ks] = as_v8<S>(l[ks][0], l[ks][1], l[ks][2], l[ks][3]);
        }
    }
    f32x16 acc[G], cor[G];
#pragma unroll
    for (int q = 0; q < G; ++q) {
        acc[q] = (f32x16){};
        cor[q] = (f32x16){};
    }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
#pragma unroll
        for (int q = 0; q < G; ++q) {
            acc[q] = M::mma(g.hi[ks], bh[q][ks], acc[q]);
            cor[q] = M::mma(g.hi[ks], bl[q][ks], cor[q]);
        }
#pragma unroll
        for (int q = 0; q < G; ++q) cor[q] = M::mma(g.lo[ks], bh[q][ks], cor[q]);
    }
#pragma unroll
    for (int q = 0; q < G; ++q) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            v[q][i] = {acc[q][i] + cor[q][i], acc[q][8 + i] + cor[q][8 + i]};
            if constexpr (M::kScale)
                v[q][i] = {ldexpf(v[q][i].x, ex[q]), ldexpf(v[q][i].y, ex[q])};
        }
    }
}

}  // namespace sfft
