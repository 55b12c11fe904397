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
    // column passes of two-pass plans: slot g is column c = g mod spt of
    // transform b = g / spt; element e of it sits at b bs + c ss + e es
    // (element strides of the input and of the output)
    FastDiv spt;
    int32_t in_bs, in_ss, in_es;
    int32_t out_bs, out_ss, out_es;
    int32_t post_tw, ptw_shift;  // first pass: output element e times W_N^{c e}
    const C2<T> *ptw_fine, *ptw_coarse;
    const C2<T> *ptw_full;  // or null: W_N^{c e} at c N + e (N: this pass' length)
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

// Schedule knobs of a compiled plan:
//   PFS   stage that issues the next group's loads (>= 0: that stage,
//         -1 the last, -2 the one before)
//   MINW  minimum waves per SIMD the register allocation must allow (0: the
//         compiler's choice) -- more resident blocks per CU, fewer registers
//   PAIR  load twiddles two per 16-byte load (else one per load)
//   PF_FIRST  in the prefetch stage, issue the prefetch before (1) or after
//         (0) the next stage's twiddle loads
//   PADSH one LDS pad element per 2^PADSH elements (kLdsNoPad: none)
//   NT    non-temporal HBM accesses: bit 0 the loads of the input (a single
//         pass, or a first column pass), bit 1 the stores of the final
//         output (a single pass, or a second column pass), bit 2 also in the
//         inverse real kernel of a real pair (else only its forward one);
//         the intermediate of a two-pass plan always keeps the default policy.
//         (On the real 2048 pair of the bench step, non-temporal accesses in
//         the forward kernel made it 10 % faster and the inverse leg that
//         follows it 10 % slower, with no net gain; they are not used there.)
constexpr int kLdsNoPad = 30;
template <int PFS_, int MINW_, int PAIR_ = 1, int PF_FIRST_ = 0, int PADSH_ = kLdsPadShift,
          int NT_ = 0>
struct Sched {
    static constexpr int PFS = PFS_;
    static constexpr int MINW = MINW_;
    static constexpr int PAIR = PAIR_;
    static constexpr int PF_FIRST = PF_FIRST_;
    static constexpr int PADSH = PADSH_;
    static constexpr int NT = NT_;
};

// Element type of a compiled kernel: float / double (stored as is), or
// Stored<StoreF16 / StoreBF16> (16-bit pairs in HBM, float arithmetic).
template <class S>
struct Stored {};
template <typename E>
struct FixedElem {
    using T = E;
    using S = StoreNative;
    static constexpr int prec = sizeof(E) == 8 ? 1 : 0;
};
template <class S_>
struct FixedElem<Stored<S_>> {
    using T = float;
    using S = S_;
    static constexpr int prec = std::is_same<S_, StoreF16>::value ? 2 : 3;
};

template <typename T_, int N, int TPT, int TPW, int MODE, class SCH, int... Rs>
struct Fixed {
    using T = typename FixedElem<T_>::T;
    using ST = Storage<T, typename FixedElem<T_>::S>;
    static constexpr int PFS = SCH::PFS;
    static constexpr int MINW = SCH::MINW;
    static constexpr int NST = sizeof...(Rs);
    static constexpr int RAD[NST] = {Rs...};
    static constexpr int THREADS = TPT * TPW;
    // Column modes: stage 0 (and the last stage of the second pass) map
    // adjacent lanes to adjacent slots (columns), so the strided HBM accesses
    // coalesce across columns; the other stages keep slot-major lanes.
    static constexpr bool COL = MODE == kModeColRow || MODE == kModeColCol;
    static constexpr bool COL_OUT = MODE == kModeColCol;
    static constexpr int PADSH = SCH::PADSH;
    static constexpr int pad(int e) { return e + (e >> PADSH); }
    static constexpr int STRIDE = N + (N >> PADSH) + 1;
    static constexpr uint32_t EB = sizeof(C2<T>);  // bytes per element in LDS
    static constexpr uint32_t DB =                  // and in HBM
        ST::native ? sizeof(C2<T>) : 2 * sizeof(uint16_t);
    static constexpr bool NT_MODE = MODE != kModeC2R || (SCH::NT & 4);
    static constexpr bool NT_LD = ST::native && (SCH::NT & 1) && NT_MODE && MODE != kModeColCol;
    static constexpr bool NT_ST = ST::native && (SCH::NT & 2) && NT_MODE && MODE != kModeColRow;
    using V2 = T __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ C2<T> ld(const char *p) {
        if constexpr (NT_LD) {
            const V2 v = __builtin_nontemporal_load(reinterpret_cast<const V2 *>(p));
            return {v.x, v.y};
        } else {
            return ST::load(p, 0);
        }
    }
    static __device__ __forceinline__ void st(char *p, C2<T> v) {
        if constexpr (NT_ST) {
            V2 w;
            w.x = v.x;
            w.y = v.y;
            __builtin_nontemporal_store(w, reinterpret_cast<V2 *>(p));
        } else {
            ST::store(p, 0, v);
        }
    }
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
            if (ns_of(i) > 1) o += ns_of(i) * tw_group(RAD[i]);
        return o;
    }
    static constexpr int bfly_of(int s) { return (N / RAD[s] + TPT - 1) / TPT; }
    static constexpr int PF = bfly_of(0) * RAD[0];
    // The next group's first-stage loads are issued in stage PF_STAGE, after
    // the twiddle loads of that stage and of the next one: vector-memory
    // counters drain in issue order, so waiting for a twiddle never waits for
    // the (slower) HBM prefetch.
    static constexpr int PF_STAGE =
        PFS >= 0 ? (PFS < NST ? PFS : NST - 1) : ((NST + PFS) > 0 ? (NST + PFS) : 0);
    static_assert(ns_of(NST) == N, "radices must multiply to N");
    static_assert(!COL || NST >= 2, "column kernels need two stages or more");

    // thread -> (slot, element lane) of stage S; gbase = first slot of the group
    template <int S>
    static __device__ __forceinline__ void map_of(const FixedArgs<T> &a, int gbase, int &g,
                                                  int &slot, int &lid, bool &active) {
        constexpr bool CM = COL && (S == 0 || (S == NST - 1 && COL_OUT));
        const int tid = (int)threadIdx.x;
        slot = CM ? tid % TPW : tid / TPT;
        lid = CM ? tid / TPW : tid % TPT;
        g = gbase + slot;
        active = g < a.total_slots;
    }

    // twiddles of stage S for this thread's butterflies
    template <int S, bool = (S < NST)>
    struct TwBuf {
        C2<T> w[1][2];
    };
    template <int S>
    struct TwBuf<S, true> {
        C2<T> w[bfly_of(S)][RAD[S] > 1 ? RAD[S] : 2];
    };

    template <int S>
    static __device__ __forceinline__ void load_tw(const FixedArgs<T> &a, TwBuf<S> &tb,
                                                   bool active, int lid) {
        if constexpr (S < NST) {
            constexpr int R = RAD[S], NS = ns_of(S), NB = N / R, B = bfly_of(S);
            constexpr int TWG = tw_group(R);
            if constexpr (NS > 1) {
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = lid + b * TPT;
                    if (active && ((NB % TPT) == 0 || j < NB)) {
                        const int k = j % NS;
                        const C2<T> *wp = a.tw + (tw_of(S) + k * TWG);
                        if constexpr (sizeof(T) == 4 && SCH::PAIR && (TWG % 2) == 0) {
                            // 16-byte aligned groups: two twiddles per load
                            const float4 *w4 = reinterpret_cast<const float4 *>(wp);
#pragma unroll
                            for (int p = 0; p < TWG / 2; ++p) {
                                const float4 q = w4[p];
                                if (2 * p + 1 < R) tb.w[b][2 * p + 1] = {q.x, q.y};
                                if (2 * p + 2 < R) tb.w[b][2 * p + 2] = {q.z, q.w};
                            }
                        } else {
#pragma unroll
                            for (int r = 1; r < R; ++r) tb.w[b][r] = wp[r - 1];
                        }
                    }
                }
            }
        }
    }

    // stage-0 element loads of slot g into pf (contiguous batch, or a column)
    static __device__ __forceinline__ void load0(const FixedArgs<T> &a, C2<T> (&pf)[PF], int g,
                                                 int lid) {
        constexpr int R = RAD[0], NB = N / R, B = bfly_of(0);
        if (g >= a.total_slots) return;
        const char *src = reinterpret_cast<const char *>(a.in);
        if constexpr (COL) {
            const uint32_t t = fdiv((uint32_t)g, a.spt);
            const uint32_t c = (uint32_t)g - t * a.spt.d;
            const uint32_t es = (uint32_t)a.in_es * DB;
            const uint32_t lane =
                (t * (uint32_t)a.in_bs + c * (uint32_t)a.in_ss) * DB + (uint32_t)lid * es;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                if ((NB % TPT) == 0 || lid + b * TPT < NB) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        pf[b * R + r] = ld(src + (lane + (uint32_t)(b * TPT + r * NB) * es));
                }
            }
            return;
        }
        const uint32_t lane = ((uint32_t)g * (uint32_t)N + (uint32_t)lid) * DB;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            if ((NB % TPT) == 0 || lid + b * TPT < NB) {
#pragma unroll
                for (int r = 0; r < R; ++r)
                    pf[b * R + r] = ld(src + (lane + (uint32_t)((b * TPT + r * NB) * DB)));
            }
        }
    }

    // Stage S.  TW_READY: its twiddles were loaded by the previous stage.
    // PF_HERE: also load stage S+1's twiddles into tn, then issue the next
    // group's first-stage loads into pf.
    template <int S, bool TW_READY, bool PF_HERE>
    static __device__ __forceinline__ void stage(const FixedArgs<T> &a, char *lds_c, int g_in,
                                                 bool active_in, int slot_in, int lid_in,
                                                 C2<T> (&pf)[PF], int g_next, TwBuf<S> &tb,
                                                 TwBuf<S + 1> &tn) {
        constexpr int R = RAD[S], NS = ns_of(S), NB = N / R, B = bfly_of(S);
        // column modes: g_in is the group's first slot and each stage maps
        // its own lanes (the next stage's twiddles and the prefetch use theirs)
        int g = g_in, slot = slot_in, lid = lid_in, lid_n = lid_in, lid_pf = lid_in;
        bool active = active_in, active_n = active_in;
        if constexpr (COL) {
            map_of<S>(a, g_in, g, slot, lid, active);
            int gx, sx;
            map_of<S + 1>(a, g_in, gx, sx, lid_n, active_n);
            bool ax;
            map_of<0>(a, g_in, gx, sx, lid_pf, ax);
        }
        constexpr bool FIRST = (S == 0), LAST = (S == NST - 1);
        constexpr bool SRC_PF = FIRST && MODE != kModeC2R;
        constexpr bool DST_G = LAST && MODE != kModeR2C;
        constexpr bool SEP_R = ct_pow2(NB) || (NB % (1 << PADSH)) == 0;
        constexpr bool SEP_W = (NS % (1 << PADSH)) == 0 || (ct_pow2(NS) && ct_pow2(R));
        const uint32_t lsb = (uint32_t)(slot * STRIDE) * EB;

        auto jvalid = [&](int b) { return (NB % TPT) == 0 || lid + b * TPT < NB; };

        C2<T> v[B][R];
        if constexpr (SRC_PF) {
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    v[b][r] = pf[b * R + r];
                    flip_im(v[b][r], a.flip_in);
                }
        }
        if constexpr (!TW_READY) load_tw<S>(a, tb, active, lid);
        if constexpr (PF_HERE) {
            if constexpr (SCH::PF_FIRST) {
                if (MODE != kModeC2R && g_next >= 0) load0(a, pf, g_next, lid_pf);
                load_tw<S + 1>(a, tn, active_n, lid_n);
            } else {
                load_tw<S + 1>(a, tn, active_n, lid_n);
                if (MODE != kModeC2R && g_next >= 0) load0(a, pf, g_next, lid_pf);
            }
        }
        if constexpr (SRC_PF) {
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
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    if constexpr (SEP_R) {
                        const uint32_t la = lsb + (uint32_t)pad(j) * EB;
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            v[b][r] = *reinterpret_cast<const C2<T> *>(
                                lds_c + (la + (uint32_t)(pad(r * NB) * EB)));
                    } else {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            v[b][r] = *reinterpret_cast<const C2<T> *>(
                                lds_c + (lsb + (uint32_t)pad(j + r * NB) * EB));
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
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], tb.w[b][r]);
                }
                dft<R, T>(v[b]);
            }
        }
        if constexpr (DST_G && COL) {
            char *dst = reinterpret_cast<char *>(a.out);
            const uint32_t t = fdiv((uint32_t)g, a.spt);
            const uint32_t c = (uint32_t)g - t * a.spt.d;
            const uint32_t es = (uint32_t)a.out_es * DB;
            const uint32_t sbase = (t * (uint32_t)a.out_bs + c * (uint32_t)a.out_ss) * DB;
            const uint32_t mask = (1u << a.ptw_shift) - 1u;
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int t2 = 0; t2 < R; ++t2) flip_im(v[b][t2], a.flip_out);
            if (a.post_tw && a.ptw_full) {
                // the four-step factors from the full table: the lanes of a
                // slot read consecutive entries (lanes without a butterfly
                // read entry 0; never stored)
                const uint32_t crow = active ? c * (uint32_t)N : 0u;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = (active && jvalid(b)) ? lid + b * TPT : 0;
                    const int o = (j / NS) * NS * R + j % NS;
#pragma unroll
                    for (int t2 = 0; t2 < R; ++t2)
                        v[b][t2] = cmul(v[b][t2], a.ptw_full[crow + (uint32_t)(o + t2 * NS)]);
                }
            } else if (a.post_tw) {
                // the four-step factors of every element, loaded together
                // (lanes without a butterfly read entry 0; never stored)
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = (active && jvalid(b)) ? lid + b * TPT : 0;
                    const int o = (j / NS) * NS * R + j % NS;
#pragma unroll
                    for (int t2 = 0; t2 < R; ++t2) {
                        const uint32_t m = active ? c * (uint32_t)(o + t2 * NS) : 0u;
                        v[b][t2] = cmul(v[b][t2],
                                        cmul(a.ptw_fine[m & mask], a.ptw_coarse[m >> a.ptw_shift]));
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    const int o = (j / NS) * NS * R + j % NS;
#pragma unroll
                    for (int t2 = 0; t2 < R; ++t2)
                        st(dst + (sbase + (uint32_t)(o + t2 * NS) * es), v[b][t2]);
                }
            }
        } else if constexpr (DST_G) {
            char *dst = reinterpret_cast<char *>(a.out);
            const uint32_t sbase = (uint32_t)g * (uint32_t)N;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * TPT;
                if (active && jvalid(b)) {
                    const int o = (j / NS) * NS * R + j % NS;
                    const uint32_t lo = (sbase + (uint32_t)o) * DB;
#pragma unroll
                    for (int t = 0; t < R; ++t) {
                        C2<T> y = v[b][t];
                        flip_im(y, a.flip_out);
                        if (MODE == kModeC2R && a.half_shift) y = cmulc(y, a.rot[o + t * NS]);
                        st(dst + (lo + (uint32_t)(t * NS * DB)), y);
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
                        const uint32_t la = lsb + (uint32_t)pad(o) * EB;
#pragma unroll
                        for (int t = 0; t < R; ++t)
                            *reinterpret_cast<C2<T> *>(
                                lds_c + (la + (uint32_t)(pad(t * NS) * EB))) = v[b][t];
                    } else {
#pragma unroll
                        for (int t = 0; t < R; ++t)
                            *reinterpret_cast<C2<T> *>(
                                lds_c + (lsb + (uint32_t)pad(o + t * NS) * EB)) = v[b][t];
                    }
                }
            }
            __syncthreads();
        }
    }

    template <int S, bool TW_READY>
    static __device__ __forceinline__ void stages_from(const FixedArgs<T> &a, char *lds_c, int g,
                                                       bool active, int slot, int lid,
                                                       C2<T> (&pf)[PF], int g_next,
                                                       TwBuf<S> &tb) {
        if constexpr (S < NST) {
            TwBuf<S + 1> tn;
            constexpr bool PF_HERE = (S == PF_STAGE);
            stage<S, TW_READY, PF_HERE>(a, lds_c, g, active, slot, lid, pf, g_next, tb, tn);
            stages_from<S + 1, PF_HERE>(a, lds_c, g, active, slot, lid, pf, g_next, tn);
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
            // kRealBatch pairs at a time, their loads issued before any use
            for (int i0 = (mod ? 0 : 1) + lid; i0 <= H / 2; i0 += kRealBatch * TPT) {
                C2<T> va[kRealBatch], vb[kRealBatch], tw[kRealBatch];
#pragma unroll
                for (int k = 0; k < kRealBatch; ++k) {
                    const int i = i0 + k * TPT;
                    if (i <= H / 2) {
                        const int ci = mod ? (H - 1 - i) : (H - i);
                        va[k] = x[i];
                        vb[k] = x[ci];
                        tw[k] = a.real_tw[i];
                    }
                }
#pragma unroll
                for (int k = 0; k < kRealBatch; ++k) {
                    const int i = i0 + k * TPT;
                    if (i > H / 2 || real_pair_skipped(i, H, mod)) continue;
                    const int ci = mod ? (H - 1 - i) : (H - i);
                    const C2<T> v = va[k], v2 = vb[k];
                    const C2<T> odd = {v.x + v2.x, v.y - v2.y};
                    const C2<T> erot = {v.x - v2.x, v.y + v2.y};
                    const C2<T> ev = cmulc(erot, tw[k]);
                    const C2<T> p = cadd(odd, ev);
                    const C2<T> m = csub(odd, ev);
                    ls[pad(i)] = cconj(p);
                    ls[pad(ci)] = m;
                }
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
            // (one pair at a time: grouping the pairs' loads measured slower
            // here, 4.15 -> 3.91 TB/s at 2048, unlike the pre-process)
            for (int i = (mod ? 0 : 1) + lid; i <= H / 2; i += TPT) {
                if (real_pair_skipped(i, H, mod)) continue;
                const int ci = mod ? (H - 1 - i) : (H - i);
                const C2<T> zi = ls[pad(i)], zc = ls[pad(ci)];
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
        // stage-0 lanes (the prefetch): column-major in the column modes
        const int slot0 = COL ? tid % TPW : slot, lid0 = COL ? tid / TPW : lid;
        const int ngroups = (a.total_slots + TPW - 1) / TPW;
        C2<T> pf[PF];
        if constexpr (COL) {
            // Column tiles share HBM lines with their neighbours when a tile
            // is narrower than 128 bytes, or when the rows of a strided side
            // do not start on 128-byte lines (a row pitch that is no multiple
            // of 128 bytes): then each window of gridDim.x groups is numbered
            // so that every XCD (blocks go round-robin over them) runs a
            // contiguous run of groups, whose shared lines meet in one L2.
            const int G = (int)gridDim.x;
            constexpr bool NARROW = TPW * (int)DB < 128;
            const auto odd_pitch = [](int32_t es) {
                return es != 1 && (((uint32_t)es * DB) & 127u) != 0;
            };
            const bool share = NARROW || odd_pitch(a.in_es) || odd_pitch(a.out_es);
            const int off = share ? (int)xcd_tile(blockIdx.x, (unsigned)G) : (int)blockIdx.x;
            int grp = off;
            if (grp < ngroups) load0(a, pf, grp * TPW + slot0, lid0);
            for (int w = 0; w < ngroups; w += G, grp += G) {
                if (grp >= ngroups) break;  // the last window's spare blocks
                const int gn = grp + G;
                const int g_next = gn < ngroups ? gn * TPW + slot0 : -1;
                TwBuf<0> t0;
                stages_from<0, false>(a, lds_c, grp * TPW, grp * TPW < a.total_slots, slot, lid,
                                      pf, g_next, t0);
            }
            return;
        }
        int grp = (int)blockIdx.x;
        if (MODE != kModeC2R && grp < ngroups) load0(a, pf, grp * TPW + slot0, lid0);
        for (; grp < ngroups; grp += (int)gridDim.x) {
            // column modes pass the group's first slot; the stages map lanes
            const int g = grp * TPW + (COL ? 0 : slot);
            const bool active = g < a.total_slots;
            const int gn = grp + (int)gridDim.x;
            const int g_next = gn < ngroups ? gn * TPW + slot0 : -1;
            if constexpr (MODE == kModeC2R) real_pre(a, lds_c, g, active, slot, lid);
            TwBuf<0> t0;
            stages_from<0, false>(a, lds_c, g, active, slot, lid, pf, g_next, t0);
            if constexpr (MODE == kModeR2C) real_post(a, lds_c, g, active, slot, lid);
        }
    }
};

template <class F>
__global__ void __launch_bounds__(F::THREADS, F::MINW > 0 ? F::MINW : 1)
    k_fixed(FixedArgs<typename F::T> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    F::run(a, reinterpret_cast<char *>(smem_raw));
}

// Registry entry of a compiled fixed kernel.
struct FixedKernel {
    int n;          // complex length of the core transform
    int variant;    // schedule variant (0 = default)
    int prec;       // 0 F32, 1 F64
    int real;       // 0: C2C kernel; 1: real pair (R2C forward + C2R inverse);
                    // 2: column passes of two-pass plans (kModeColRow, kModeColCol)
    int tpt, tpw;
    int nst;
    int radices[kMaxStages];
    size_t lds_bytes;
    // launch(args, mode, grid, stream): mode kModeC2C / kModeR2C / kModeC2R
    hipError_t (*launch)(const void *args, int mode, unsigned grid, hipStream_t stream);
    int (*blocks_per_cu)(int mode);
};

void register_fixed(const FixedKernel &k);
// variant < 0: the default (variant 0); returns nullptr if none
const FixedKernel *find_fixed(int n, int prec, int real, int variant = -1);
int fixed_variant_count(int n, int prec, int real);

template <class F>
hipError_t fixed_launch_one(const void *args, unsigned grid, hipStream_t stream) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_fixed<F>) != hipSuccess) return lds_err;
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
template <typename T, int N, int TPT, int TPW, class SCH, int VARIANT, int... Rs>
struct FixedC2C {
    using F = Fixed<T, N, TPT, TPW, kModeC2C, SCH, Rs...>;
    FixedC2C() {
        FixedKernel k{};
        k.n = N;
        k.variant = VARIANT;
        k.prec = FixedElem<T>::prec;
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
template <typename T, int N, int TPT, int TPW, class SCH, int VARIANT, int... Rs>
struct FixedReal {
    using FR = Fixed<T, N, TPT, TPW, kModeR2C, SCH, Rs...>;
    using FI = Fixed<T, N, TPT, TPW, kModeC2R, SCH, Rs...>;
    FixedReal() {
        FixedKernel k{};
        k.n = N;
        k.variant = VARIANT;
        k.prec = FixedElem<T>::prec;
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

// Registers the column-pass pair of one compile-time plan of column length
// N (registry kind 2): the first pass of a two-pass plan (kModeColRow) and
// the second (kModeColCol).
template <typename T, int N, int TPT, int TPW, class SCH, int VARIANT, int... Rs>
struct FixedCol {
    using F1 = Fixed<T, N, TPT, TPW, kModeColRow, SCH, Rs...>;
    using F2 = Fixed<T, N, TPT, TPW, kModeColCol, SCH, Rs...>;
    FixedCol() {
        FixedKernel k{};
        k.n = N;
        k.variant = VARIANT;
        k.prec = FixedElem<T>::prec;
        k.real = 2;
        k.tpt = TPT;
        k.tpw = TPW;
        fixed_fill_radices<Rs...>(k);
        k.lds_bytes = F1::LDS_BYTES;
        k.launch = [](const void *args, int mode, unsigned grid, hipStream_t s) {
            return mode == kModeColCol ? fixed_launch_one<F2>(args, grid, s)
                                       : fixed_launch_one<F1>(args, grid, s);
        };
        k.blocks_per_cu = [](int mode) {
            return mode == kModeColCol ? fixed_occupancy<F2>() : fixed_occupancy<F1>();
        };
        register_fixed(k);
    }
};

}  // namespace sfft
