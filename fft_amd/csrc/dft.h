// dft.h -- register-resident small DFTs for the Stockham stages (gfx950).
//
// Every DFT here is the forward transform y[k] = sum_n x[n] W_R^{nk},
// W_R = e^{-2 pi i / R}, in natural order in and out.  The inverse transform
// of the whole FFT is taken as conj(FFT(conj(x))), so no inverse butterflies
// exist.  Radices:
//   base:       1, 2, 3, 4, 5 (hand-scheduled), any odd R (symmetric pairs)
//   composite:  R = A * B from two smaller DFTs and compile-time twiddles
// All indices are compile-time constants, so the arrays live in VGPRs.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace sfft {

template <typename T>
struct alignas(2 * sizeof(T)) C2 {
    T x, y;
};

// Fused multiply-add spelled out: the library builds with -ffp-contract=off,
// so every kernel that runs the same radices computes bit-identical results.
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) {
    return {fmadd(a.x, b.x, -(a.y * b.y)), fmadd(a.x, b.y, a.y * b.x)};
}
// a * conj(b)
template <typename T>
__device__ __forceinline__ C2<T> cmulc(C2<T> a, C2<T> b) {
    return {fmadd(a.x, b.x, a.y * b.y), fmadd(a.y, b.x, -(a.x * b.y))};
}
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__device__ __forceinline__ C2<T> cconj(C2<T> a) { return {a.x, -a.y}; }
// -i * a
template <typename T>
__device__ __forceinline__ C2<T> cmul_negi(C2<T> a) { return {a.y, -a.x}; }
// +i * a
template <typename T>
__device__ __forceinline__ C2<T> cmul_posi(C2<T> a) { return {-a.y, a.x}; }

// ---------------------------------------------------------------------------
// Compile-time trigonometry of rational angles 2 pi p / q.
// Octant reduction on the exact rational, then a Taylor series on
// |angle| <= pi/4 (17 terms: far below double rounding).
// ---------------------------------------------------------------------------
namespace ct {
constexpr double kPi = 3.141592653589793238462643383279502884;

constexpr double taylor_sin(double x) {
    double t = x, s = x;
    for (int k = 1; k < 17; ++k) {
        t *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        s += t;
    }
    return s;
}
constexpr double taylor_cos(double x) {
    double t = 1.0, s = 1.0;
    for (int k = 1; k < 17; ++k) {
        t *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
        s += t;
    }
    return s;
}

// cos(2 pi p / q) and sin(2 pi p / q) for integers, q > 0.
struct CS {
    double c, s;
};
constexpr CS cossin(long long p, long long q) {
    p %= q;
    if (p < 0) p += q;
    // fraction f = p / q in [0, 1).  Work with f scaled by 8q to stay integral.
    long long num = 8 * p, den = 8 * q;  // f = num / den
    double sign_s = 1.0, sign_c = 1.0;
    if (2 * num > den) {  // f > 1/2: sin odd about 1/2, cos even
        num = den - num;
        sign_s = -1.0;
    }
    // f in [0, 1/2]
    if (4 * num > den) {  // f > 1/4: cos(2pi f) = -cos(2pi(1/2 - f)), sin same
        num = den / 2 - num;
        sign_c = -sign_c;
    }
    // f in [0, 1/4]
    bool swap = false;
    if (8 * num > den) {  // f > 1/8: swap cos/sin with 1/4 - f
        num = den / 4 - num;
        swap = true;
    }
    double a = 2.0 * kPi * (double)num / (double)den;
    double c = taylor_cos(a), s = taylor_sin(a);
    if (swap) {
        double t = c;
        c = s;
        s = t;
    }
    return CS{sign_c * c, sign_s * s};
}
}  // namespace ct

// x * W_R^E with W_R = e^{-2 pi i / R}; E, R compile-time.  Exact special
// cases for multiples of 1/8 of a turn, general constant multiply otherwise.
template <int E, int R, typename T>
__device__ __forceinline__ C2<T> twc(C2<T> x) {
    constexpr int e = ((E % R) + R) % R;
    if constexpr (e == 0) {
        return x;
    } else if constexpr (2 * e == R) {
        return {-x.x, -x.y};
    } else if constexpr (4 * e == R) {
        return cmul_negi(x);
    } else if constexpr (4 * e == 3 * R) {
        return cmul_posi(x);
    } else if constexpr (8 * e == R || 8 * e == 3 * R || 8 * e == 5 * R || 8 * e == 7 * R) {
        constexpr T h = (T)0.70710678118654752440084436210484904;
        // W = e^{-i pi e'/4}
        if constexpr (8 * e == R) return {(x.x + x.y) * h, (x.y - x.x) * h};        // (1-i)/sqrt2
        else if constexpr (8 * e == 3 * R) return {(x.y - x.x) * h, -(x.x + x.y) * h};  // (-1-i)/sqrt2
        else if constexpr (8 * e == 5 * R) return {-(x.x + x.y) * h, (x.x - x.y) * h};  // (-1+i)/sqrt2
        else return {(x.x - x.y) * h, (x.x + x.y) * h};                                 // (1+i)/sqrt2
    } else {
        constexpr ct::CS w = ct::cossin(e, R);
        constexpr T c = (T)w.c;
        constexpr T s = (T)(-w.s);
        return {fmadd(x.x, c, -(x.y * s)), fmadd(x.x, s, x.y * c)};
    }
}

template <int I, int N, typename F>
__device__ __forceinline__ void sfor(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        sfor<I + 1, N>(f);
    }
}

template <int R, typename T>
__device__ __forceinline__ void dft(C2<T> (&v)[R]);

// ----------------------------------------------------------------- bases --

template <typename T>
__device__ __forceinline__ void dft2(C2<T> &a, C2<T> &b) {
    C2<T> t = a;
    a = cadd(t, b);
    b = csub(t, b);
}

template <typename T>
__device__ __forceinline__ void dft_base1(C2<T> (&)[1]) {}

template <typename T>
__device__ __forceinline__ void dft_base2(C2<T> (&v)[2]) { dft2(v[0], v[1]); }

template <typename T>
__device__ __forceinline__ void dft_base3(C2<T> (&v)[3]) {
    constexpr T c = (T)-0.5;
    constexpr T s = (T)0.86602540378443864676372317075293618;
    C2<T> t1 = cadd(v[1], v[2]);
    C2<T> t2 = {fmadd(c, t1.x, v[0].x), fmadd(c, t1.y, v[0].y)};
    C2<T> t3 = {s * (v[1].x - v[2].x), s * (v[1].y - v[2].y)};
    v[0] = cadd(v[0], t1);
    v[1] = {t2.x + t3.y, t2.y - t3.x};  // t2 - i t3
    v[2] = {t2.x - t3.y, t2.y + t3.x};  // t2 + i t3
}

template <typename T>
__device__ __forceinline__ void dft_base4(C2<T> (&v)[4]) {
    C2<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C2<T> s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = {d02.x + d13.y, d02.y - d13.x};  // d02 - i d13
    v[3] = {d02.x - d13.y, d02.y + d13.x};  // d02 + i d13
}

// Odd R: symmetric pairs a_m = x_m + x_{R-m}, b_m = x_m - x_{R-m};
// y_k = x_0 + sum c_mk a_m -/+ i sum s_mk b_m for k and R-k.
template <int R, typename T>
__device__ __forceinline__ void dft_odd(C2<T> (&v)[R]) {
    constexpr int H = (R - 1) / 2;
    C2<T> a[H], b[H];
    sfor<0, H>([&](auto mi) {
        constexpr int m = decltype(mi)::value + 1;
        a[m - 1] = cadd(v[m], v[R - m]);
        b[m - 1] = csub(v[m], v[R - m]);
    });
    C2<T> x0 = v[0];
    C2<T> y0 = x0;
    sfor<0, H>([&](auto mi) { y0 = cadd(y0, a[decltype(mi)::value]); });
    sfor<0, H>([&](auto ki) {
        constexpr int k = decltype(ki)::value + 1;
        C2<T> t = x0, u = {(T)0, (T)0};
        sfor<0, H>([&](auto mi) {
            constexpr int m = decltype(mi)::value + 1;
            constexpr ct::CS w = ct::cossin((long long)m * k, R);
            constexpr T c = (T)w.c, s = (T)w.s;
            t.x = fmadd(c, a[m - 1].x, t.x);
            t.y = fmadd(c, a[m - 1].y, t.y);
            u.x = fmadd(s, b[m - 1].x, u.x);
            u.y = fmadd(s, b[m - 1].y, u.y);
        });
        v[k] = {t.x + u.y, t.y - u.x};      // t - i u
        v[R - k] = {t.x - u.y, t.y + u.x};  // t + i u
    });
    v[0] = y0;
}

// ------------------------------------------------------------- composite --
// R = A * B, n = a + A b, k = k2 + B k1:
//   z[a][k2] = W_R^{a k2} DFT_B(x[a + A b])[k2];  y[k2 + B k1] = DFT_A(z[.][k2])[k1]
template <int A, int B, typename T>
__device__ __forceinline__ void dft_composite(C2<T> (&v)[A * B]) {
    constexpr int R = A * B;
    C2<T> z[A][B];
    sfor<0, A>([&](auto ai) {
        constexpr int a = decltype(ai)::value;
        C2<T> t[B];
        sfor<0, B>([&](auto bi) { t[decltype(bi)::value] = v[a + A * decltype(bi)::value]; });
        dft<B, T>(t);
        sfor<0, B>([&](auto ki) {
            constexpr int k2 = decltype(ki)::value;
            z[a][k2] = twc<a * k2, R>(t[k2]);
        });
    });
    sfor<0, B>([&](auto ki) {
        constexpr int k2 = decltype(ki)::value;
        C2<T> t[A];
        sfor<0, A>([&](auto ai) { t[decltype(ai)::value] = z[decltype(ai)::value][k2]; });
        dft<A, T>(t);
        sfor<0, A>([&](auto ji) {
            constexpr int k1 = decltype(ji)::value;
            v[k2 + B * k1] = t[k1];
        });
    });
}

template <int R, typename T>
__device__ __forceinline__ void dft(C2<T> (&v)[R]) {
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft_base2(v);
    } else if constexpr (R == 3) {
        dft_base3(v);
    } else if constexpr (R == 4) {
        dft_base4(v);
    } else if constexpr (R == 8) {
        dft_composite<2, 4>(v);
    } else if constexpr (R == 16) {
        dft_composite<4, 4>(v);
    } else if constexpr (R == 32) {
        dft_composite<4, 8>(v);
    } else if constexpr (R == 6) {
        dft_composite<2, 3>(v);
    } else if constexpr (R == 10) {
        dft_composite<2, 5>(v);
    } else if constexpr (R == 12) {
        dft_composite<4, 3>(v);
    } else if constexpr (R == 15) {
        dft_composite<3, 5>(v);
    } else if constexpr (R == 20) {
        dft_composite<4, 5>(v);
    } else if constexpr (R == 9) {
        dft_composite<3, 3>(v);
    } else if constexpr (R == 25) {
        dft_composite<5, 5>(v);
    } else if constexpr (R % 2 == 1) {
        dft_odd<R>(v);
    } else {
        static_assert(R % 2 == 1, "unsupported radix");
    }
}

}  // namespace sfft
