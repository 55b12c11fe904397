// stockham.h -- LDS-resident Stockham block FFT for gfx950.
//
// One workgroup holds `tpw` slots (sub-transforms) of length n in LDS.  The
// first stage reads HBM directly (coalesced: consecutive lanes take
// consecutive elements, or consecutive slots for column-strided input), every
// stage boundary is one LDS exchange, and the last stage writes HBM directly.
// The inverse transform is conj(FFT(conj(x))): conj on load and on store.
#pragma once

#include <mutex>

#include "dft.h"
#include "plan_desc.h"
#include "storage.h"

namespace sfft {

// Raises a kernel's dynamic LDS limit to what one CU holds, once per kernel
// and process: the first caller sets the attribute, every caller (on any
// thread) gets that call's result.
inline hipError_t allow_max_lds(std::once_flag &once, hipError_t &err, const void *kernel) {
    std::call_once(once, [&] {
        err = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024 - kLdsReserve);
        if (err != hipSuccess) (void)hipGetLastError();
    });
    return err;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
    if (f.d <= 1) return n;
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

__device__ __forceinline__ int lds_idx(int e) { return e + (e >> kLdsPadShift); }

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2).  Tile index of block b so that every XCD runs a contiguous range of
// tiles: neighbouring column tiles, which share HBM lines when a tile is
// narrower than a line, then meet in the same L2.  A bijection on [0, nb).
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb) {
    const unsigned q = nb / kXcds, r = nb % kXcds;
    const unsigned x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
}

// Pairs of the runtime real pre-/post-process whose loads are issued
// together (the loop trip count is not a compile-time constant, so the
// compiler would otherwise wait for each pair's loads before the next).
constexpr int kRealBatch = 4;

// Real-FFT fusion modes.
enum : int32_t {
    kModeC2C = 0,
    kModeR2C = 1,  // forward real: post-process pairs (i, h-i) after the last stage
    kModeC2R = 2,  // inverse real: pre-process pairs before the first stage
    // two-pass real transforms (global element index m within the transform):
    kModeR2CLoad = 3,   // first pass: half-shift rotation at the HBM load
    kModeC2RLoad = 4,   // first pass: pre-process from in[i], in[h-i] at the HBM load
    kModeC2RStore = 5,  // second pass: half-shift rotation at the HBM store
    // compiled column-pass kernels of two-pass plans (fixed.h only)
    kModeColRow = 6,  // first pass: column loads, four-step twiddle, row stores
    kModeColCol = 7,  // second pass: column loads and column stores
};

// The pair loops of the real pre-/post-process (signalsmith-fft.h:463-472,
// :482-492) visit i = 0..h/2 (1..h/2 without the half shift) with partner
// ci = h-1-i (h-i).  With the half shift and h even the pairs i = h/2-1 and
// h/2 are the same two bins, and the reference keeps the values of the later
// one: that pair is skipped so every bin has exactly one writer.
__device__ __forceinline__ bool real_pair_skipped(int i, int h, bool mod) {
    return mod && (h % 2 == 0) && i == h / 2 - 1;
}

template <typename T>
struct PassArgs {
    PassDesc pd;
    IoDesc iin, iout;
    const C2<T> *in;
    C2<T> *out;
    const C2<T> *tw;       // stage twiddles
    int32_t total_slots;   // slots in the whole launch
    int32_t conj_in;       // conjugate at the HBM load
    int32_t conj_out;      // conjugate at the HBM store
    int32_t mode;          // kModeC2C / kModeR2C / kModeC2R
    int32_t half_shift;    // ModifiedRealFFT
    int32_t post_tw;       // multiply output element e of slot g by W_N^{(g mod per_batch) e}
    int32_t post_tw_shift; // fine table covers 2^shift entries
    const C2<T> *real_tw;  // twiddlesMinusI: -i e^{-2 pi i (i [+1/2]) / N}, N/4 + 1 entries
    const C2<T> *rot;      // half-shift pre-rotations e^{-2 pi i i / N}, N/2 entries
    const C2<T> *ptw_fine;    // W_N^{m}, m < 2^shift (FP64 values rounded once to T)
    const C2<T> *ptw_coarse;  // W_N^{m 2^shift}
    const C2<T> *ptw_full;    // or null: W_N^{c e} at c n + e for every column c and
                              // element e (FP64 values rounded once to T)
    int32_t in_from_out;   // second pass of a two-pass plan: reads the output buffer
    int32_t real_h;        // complex length h of a two-pass real transform
};

// Value of pre-processed bin m (before the inverse complex FFT) from the
// half spectrum x of length h (signalsmith-fft.h:478-492), for one element.
template <typename T>
__device__ __forceinline__ C2<T> real_pre_bin(const C2<T> *x, const C2<T> *tmi, int m, int h,
                                             bool mod) {
    if (!mod && m == 0) {
        const C2<T> v = x[0];
        return {v.x + v.y, v.x - v.y};
    }
    int i, ci;
    if (mod && (h % 2 == 0) && (m == h / 2 - 1 || m == h / 2)) {
        i = h / 2;
        ci = h / 2 - 1;
    } else {
        const int mirror = mod ? (h - 1 - m) : (h - m);
        i = m < mirror ? m : mirror;
        ci = mod ? (h - 1 - i) : (h - i);
    }
    const C2<T> v = x[i], v2 = x[ci];
    const C2<T> odd = {v.x + v2.x, v.y - v2.y};
    const C2<T> erot = {v.x - v2.x, v.y + v2.y};
    const C2<T> ev = cmulc(erot, tmi[i]);
    if (m == i && i != ci) return cadd(odd, ev);
    return cconj(csub(odd, ev));
}

// The pass descriptor `a` may live in kernel arguments or in a device table
// (persistent launches); in, out and total_slots are per launch / per item.
// Optional gather lists (persistent launches) give each slot its own input /
// output offset in elements instead of the strided address map.
// S is the HBM storage format of in / out (storage.h); with a 16-bit format
// the pointers address 4-byte values and only complex-to-complex modes run.
template <typename T, int EPT, typename S = StoreNative, bool MFMA = false>
struct Block {
    using St = Storage<T, S>;
    const PassArgs<T> &a;
    const C2<T> *in;
    C2<T> *out;
    int total_slots;
    C2<T> *lds;
    int tid;
    int slot0;
    const int64_t *gather_in = nullptr;
    const int64_t *gather_out = nullptr;
    // Uniform descriptor fields, read once: the descriptor may live in global
    // memory (persistent launches), where reading it inside the element loops
    // would be repeated after every store the compiler cannot disambiguate.
    int n_, tpt_, tpw_, lds_stride_, nstages_, in_col_, out_col_;
    int conj_in_, conj_out_, mode_, half_shift_, post_tw_, post_tw_shift_, real_h_, mfma_;
    IoDesc iin_, iout_;
    const C2<T> *tw_, *real_tw_, *rot_;
    const C2<T> *ptw_fine_, *ptw_coarse_, *ptw_full_;

    __device__ Block(const PassArgs<T> &args, const C2<T> *in_, C2<T> *out_, int total, C2<T> *l,
                     int t, int s0)
        : a(args), in(in_), out(out_), total_slots(total), lds(l), tid(t), slot0(s0) {
        n_ = args.pd.n;
        tpt_ = args.pd.tpt;
        tpw_ = args.pd.tpw;
        lds_stride_ = args.pd.lds_stride;
        nstages_ = args.pd.nstages;
        in_col_ = args.pd.in_col;
        out_col_ = args.pd.out_col;
        conj_in_ = args.conj_in;
        conj_out_ = args.conj_out;
        mode_ = args.mode;
        half_shift_ = args.half_shift;
        post_tw_ = args.post_tw;
        post_tw_shift_ = args.post_tw_shift;
        real_h_ = args.real_h;
        mfma_ = args.pd.mfma;
        iin_ = args.iin;
        iout_ = args.iout;
        tw_ = args.tw;
        real_tw_ = args.real_tw;
        rot_ = args.rot;
        ptw_fine_ = args.ptw_fine;
        ptw_coarse_ = args.ptw_coarse;
        ptw_full_ = args.ptw_full;
    }

    __device__ __forceinline__ int64_t ibase(int g) const {
        return gather_in ? gather_in[g] : gbase(iin_, g);
    }
    __device__ __forceinline__ int64_t obase(int g) const {
        return gather_out ? gather_out[g] : gbase(iout_, g);
    }
    // byte address of element `elem` of an HBM buffer
    __device__ __forceinline__ static const char *slot_ptr(const void *p, int64_t elem) {
        return static_cast<const char *>(p) + elem * St::kBytes;
    }
    __device__ __forceinline__ static char *slot_ptr(void *p, int64_t elem) {
        return static_cast<char *>(p) + elem * St::kBytes;
    }

    // transform base and in-transform offset of slot g's element 0
    __device__ __forceinline__ void split_slot(const IoDesc &io, int g, int64_t &tbase,
                                               int64_t &cofs) const {
        if (io.per_batch.d <= 1) {
            tbase = (int64_t)g * io.batch_stride;
            cofs = 0;
            return;
        }
        const uint32_t b = fdiv((uint32_t)g, io.per_batch);
        tbase = (int64_t)b * io.batch_stride;
        cofs = (int64_t)((uint32_t)g - b * io.per_batch.d) * io.slot_stride;
    }

    __device__ __forceinline__ int64_t gbase(const IoDesc &io, int g) const {
        if (io.per_batch.d <= 1) return (int64_t)g * io.batch_stride;
        uint32_t b = fdiv((uint32_t)g, io.per_batch);
        uint32_t r = (uint32_t)g - b * io.per_batch.d;
        return (int64_t)b * io.batch_stride + (int64_t)r * io.slot_stride;
    }

    __device__ __forceinline__ C2<T> post_twiddle(int g, int e) const {
        // W_N^{c e}, c = column index of slot g, as the product of two table
        // entries (fine x coarse), each an FP64 value rounded once to T.
        // (or the entry of the full table)
        const uint32_t c = (uint32_t)g - fdiv((uint32_t)g, iout_.per_batch) * iout_.per_batch.d;
        if (ptw_full_) return ptw_full_[c * (uint32_t)n_ + (uint32_t)e];
        // c < N2 and e < N1, so c e < N: no reduction needed.
        const uint32_t m = c * (uint32_t)e;
        return cmul(ptw_fine_[m & ((1u << post_tw_shift_) - 1u)], ptw_coarse_[m >> post_tw_shift_]);
    }

    // One Stockham stage of radix R.  No branch separates a load from the
    // next one: the data loads of all butterflies are issued before any value
    // is used, and the table loads (twiddles, rotations, four-step factors)
    // are unconditional, with the indices of lanes without a butterfly
    // clamped to valid entries (their values are never stored), so the
    // compiler keeps them in flight together.  (A load followed by its use
    // and a branch on the pass mode waits out one latency per element.)  The
    // arithmetic per element is that of stage_wide, so the results are
    // bitwise the same.
    template <int R>
    __device__ __forceinline__ void stage(int s, bool src_global, bool dst_global, bool col_map) {
        if constexpr (R == 16 && !St::native && MFMA) {
            if (mfma_) {
                stage16_mfma(s, src_global, dst_global, col_map);
                return;
            }
        }
        if constexpr (EPT >= 32) {
            stage_wide<R>(s, src_global, dst_global, col_map);
            return;
        }
        const StageDesc st = a.pd.st[s];
        constexpr int B = (EPT / R) > 0 ? (EPT / R) : 1;
        const int n = n_;
        const int nb = n / R;
        const int ns = st.ns;
        int slot, lid;
        if (col_map) {
            slot = tid % tpw_;
            lid = tid / tpw_;
        } else {
            slot = tid / tpt_;
            lid = tid % tpt_;
        }
        const int g = slot0 + slot;
        const bool active = (slot < tpw_) && (lid < tpt_) && (g < total_slots);
        C2<T> *ls = lds + slot * lds_stride_;
        // butterflies of this thread: j = lid + b tpt, valid when ok(b)
        auto ok = [&](int b) { return active && (lid + b * tpt_ < nb); };

        C2<T> v[B][R];
        // ---- load
        if (St::native && src_global && mode_ == kModeC2RLoad) {
            // two-pass inverse real: pre-process each element from its pair
            int64_t tb = 0, cofs = 0;
            if (active) split_slot(iin_, g, tb, cofs);
            const int64_t es = iin_.elem_stride;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (ok(b)) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int m = (int)(cofs + (int64_t)(j + r * nb) * es);
                        if constexpr (St::native)
                            v[b][r] = cconj(real_pre_bin(in + tb, real_tw_, m, real_h_,
                                                         half_shift_ != 0));
                    }
                }
            }
        } else if (src_global) {
            // the slot's first element, then 32-bit element offsets within the
            // transform (a transform spans fewer than 2^32 elements): one
            // 32-bit add and one 64-bit shift-add per element address
            const char *sp = slot_ptr(in, active ? ibase(g) : 0);
            const uint32_t es = (uint32_t)iin_.elem_stride, rstep = (uint32_t)nb * es;
            // the first pass of a two-pass plan reads its input once:
            // non-temporal loads (the second pass reads the intermediate with
            // the default policy, from the caches where it can).  Not for
            // tiles narrower than a 128-byte line: neighbouring tiles share
            // those lines, and non-temporal loads fetched them again (2^24
            // F32, 4-column tiles: 1.42 -> 1.21 TB/s).
            bool nt = false;
            if constexpr (St::native) nt = post_tw_ != 0 && tpw_ * St::kBytes >= 128;
            if (nt) {
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t j = (uint32_t)(lid + b * tpt_);
                    if (ok(b)) {
#pragma unroll
                        for (int r = 0; r < R; ++r) {
                            if constexpr (St::native)
                                v[b][r] = St::load_nt(sp, (int64_t)(j * es + (uint32_t)r * rstep));
                        }
                    }
                }
            } else {
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t j = (uint32_t)(lid + b * tpt_);
                    if (ok(b)) {
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            v[b][r] = St::load(sp, (int64_t)(j * es + (uint32_t)r * rstep));
                    }
                }
            }
            if (conj_in_) {
#pragma unroll
                for (int b = 0; b < B; ++b)
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r].y = -v[b][r].y;
            }
            if (half_shift_ && (mode_ == kModeR2C || mode_ == kModeR2CLoad)) {
                // half-shift pre-rotation by element e of the transform (one
                // pass), or by its index in the whole transform (first of two)
                int64_t tb = 0, cofs = 0;
                if (active && mode_ == kModeR2CLoad) split_slot(iin_, g, tb, cofs);
                const uint32_t rs = mode_ == kModeR2CLoad ? es : 1u;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const uint32_t j = ok(b) ? (uint32_t)(lid + b * tpt_) : 0u;
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        v[b][r] = cmul(v[b][r], rot_[(uint32_t)cofs + (j + (uint32_t)(r * nb)) * rs]);
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (ok(b)) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r] = ls[lds_idx(j + r * nb)];
                }
            }
            __syncthreads();  // every slot read before any slot is overwritten
        }
        // ---- twiddle + butterfly
        if (ns > 1) {
            // all twiddle loads of the stage, then (scheduling barrier: the
            // compiler would otherwise pair each load with its use) the
            // products
            C2<T> w[B][R > 1 ? R - 1 : 1];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = ok(b) ? lid + b * tpt_ : 0;
                const uint32_t q = fdiv((uint32_t)j, st.div);
                const uint32_t k = (uint32_t)j - q * (uint32_t)ns;
                const C2<T> *wp = tw_ + (uint32_t)(st.tw_off + k * (uint32_t)tw_group(R));
#pragma unroll
                for (int r = 1; r < R; ++r) w[b][r - 1] = wp[r - 1];
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int b = 0; b < B; ++b)
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[b][r - 1]);
        }
#pragma unroll
        for (int b = 0; b < B; ++b) dft<R, T>(v[b]);
        // ---- store
        if (dst_global) {
            char *sp = slot_ptr(out, active ? obase(g) : 0);
            const uint32_t es = (uint32_t)iout_.elem_stride, rstep = (uint32_t)ns * es;
            if (conj_out_) {
#pragma unroll
                for (int b = 0; b < B; ++b)
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r].y = -v[b][r].y;
            }
            if (half_shift_ && (mode_ == kModeC2R || mode_ == kModeC2RStore)) {
                int64_t tbo = 0, cofo = 0;
                if (active && mode_ == kModeC2RStore) split_slot(iout_, g, tbo, cofo);
                const uint32_t rs = mode_ == kModeC2RStore ? es : 1u;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = ok(b) ? lid + b * tpt_ : 0;
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int o = (int)q * ns * R + (j - (int)q * ns);
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        v[b][r] = cmulc(v[b][r], rot_[(uint32_t)cofo + (uint32_t)(o + r * ns) * rs]);
                }
            }
            if (post_tw_ && ptw_full_) {
                // four-step twiddle W_N^{c e} (c: column of slot g) from the
                // full table: lanes of a slot read consecutive entries
                const uint32_t crow =
                    active ? ((uint32_t)g - fdiv((uint32_t)g, iout_.per_batch) * iout_.per_batch.d) *
                                 (uint32_t)n_
                           : 0u;
                constexpr int G = R < 8 ? R : 8;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = ok(b) ? lid + b * tpt_ : 0;
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int o = (int)q * ns * R + (j - (int)q * ns);
#pragma unroll
                    for (int r0 = 0; r0 < R; r0 += G) {
                        C2<T> fu[G];
#pragma unroll
                        for (int i = 0; i < G && r0 + i < R; ++i)
                            fu[i] = ptw_full_[crow + (uint32_t)(o + (r0 + i) * ns)];
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int i = 0; i < G && r0 + i < R; ++i)
                            v[b][r0 + i] = cmul(v[b][r0 + i], fu[i]);
                    }
                }
            } else if (post_tw_) {
                // four-step twiddle W_N^{c e} (c: column of slot g) as the
                // product of a fine and a coarse table entry
                const uint32_t c =
                    active ? (uint32_t)g - fdiv((uint32_t)g, iout_.per_batch) * iout_.per_batch.d : 0u;
                const uint32_t mask = (1u << post_tw_shift_) - 1u;
                // loads in groups of up to 8 elements, each group issued
                // before its products (scheduling barriers as above)
                constexpr int G = R < 8 ? R : 8;
#pragma unroll
                for (int b = 0; b < B; ++b) {
                    const int j = ok(b) ? lid + b * tpt_ : 0;
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int o = (int)q * ns * R + (j - (int)q * ns);
#pragma unroll
                    for (int r0 = 0; r0 < R; r0 += G) {
                        C2<T> fi[G], co[G];
#pragma unroll
                        for (int i = 0; i < G && r0 + i < R; ++i) {
                            // c < N2 and e < N1, so c e < N: no reduction needed
                            const uint32_t m = c * (uint32_t)(o + (r0 + i) * ns);
                            fi[i] = ptw_fine_[m & mask];
                            co[i] = ptw_coarse_[m >> post_tw_shift_];
                        }
                        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                        for (int i = 0; i < G && r0 + i < R; ++i)
                            v[b][r0 + i] = cmul(v[b][r0 + i], cmul(fi[i], co[i]));
                    }
                }
            }
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (ok(b)) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const uint32_t o = q * (uint32_t)(ns * R) + ((uint32_t)j - q * (uint32_t)ns);
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        St::store(sp, (int64_t)(o * es + (uint32_t)r * rstep), v[b][r]);
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (ok(b)) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const int o = (int)q * ns * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) ls[lds_idx(o + r * ns)] = v[b][r];
                }
            }
            __syncthreads();
        }
    }

    // The stage of the 32-element instances: one element's accesses at a
    // time (the grouped form of stage() needs more than the 128 registers a
    // 1024-thread block allows at 32 elements per thread).  Same arithmetic.
    template <int R>
    __device__ __forceinline__ void stage_wide(int s, bool src_global, bool dst_global,
                                               bool col_map) {
        const StageDesc st = a.pd.st[s];
        constexpr int B = (EPT / R) > 0 ? (EPT / R) : 1;
        const int n = n_;
        const int nb = n / R;
        int slot, lid;
        if (col_map) {
            slot = tid % tpw_;
            lid = tid / tpw_;
        } else {
            slot = tid / tpt_;
            lid = tid % tpt_;
        }
        const int g = slot0 + slot;
        const bool active = (slot < tpw_) && (lid < tpt_) && (g < total_slots);
        C2<T> *ls = lds + slot * lds_stride_;

        C2<T> v[B][R];
        // ---- load
        if (St::native && src_global && mode_ == kModeC2RLoad) {
            // two-pass inverse real: pre-process each element from its pair
            int64_t tb = 0, cofs = 0;
            if (active) split_slot(iin_, g, tb, cofs);
            const int64_t es = iin_.elem_stride;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (active && j < nb) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int m = (int)(cofs + (int64_t)(j + r * nb) * es);
                        if constexpr (St::native)
                            v[b][r] = cconj(real_pre_bin(in + tb, real_tw_, m, real_h_,
                                                         half_shift_ != 0));
                    }
                }
            }
        } else if (src_global) {
            const int64_t base = active ? ibase(g) : 0;
            const int64_t es = iin_.elem_stride;
            int64_t tb = 0, cofs = 0;
            if (active && mode_ == kModeR2CLoad) split_slot(iin_, g, tb, cofs);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (active && j < nb) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int e = j + r * nb;
                        C2<T> x = St::load(in, base + (int64_t)e * es);
                        if (conj_in_) x.y = -x.y;
                        if (mode_ == kModeR2C && half_shift_) x = cmul(x, rot_[e]);
                        if (mode_ == kModeR2CLoad && half_shift_)
                            x = cmul(x, rot_[cofs + (int64_t)e * es]);
                        v[b][r] = x;
                    }
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (active && j < nb) {
#pragma unroll
                    for (int r = 0; r < R; ++r) v[b][r] = ls[lds_idx(j + r * nb)];
                }
            }
            __syncthreads();  // every slot read before any slot is overwritten
        }
        // ---- twiddle + butterfly
        const int ns = st.ns;
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = lid + b * tpt_;
            if (active && j < nb) {
                if (ns > 1) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const C2<T> *w = tw_ + st.tw_off + k * tw_group(R);
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r - 1]);
                }
                dft<R, T>(v[b]);
            }
        }
        // ---- store
        if (dst_global) {
            const int64_t base = active ? obase(g) : 0;
            const int64_t es = iout_.elem_stride;
            int64_t tbo = 0, cofo = 0;
            if (active && mode_ == kModeC2RStore) split_slot(iout_, g, tbo, cofo);
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (active && j < nb) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const int o = (int)q * ns * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int e = o + r * ns;
                        C2<T> y = v[b][r];
                        if (conj_out_) y.y = -y.y;
                        if (mode_ == kModeC2R && half_shift_) y = cmulc(y, rot_[e]);
                        if (mode_ == kModeC2RStore && half_shift_)
                            y = cmulc(y, rot_[cofo + (int64_t)e * es]);
                        if (post_tw_) y = cmul(y, post_twiddle(g, e));
                        St::store(out, base + (int64_t)e * es, y);
                    }
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * tpt_;
                if (active && j < nb) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const int o = (int)q * ns * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) ls[lds_idx(o + r * ns)] = v[b][r];
                }
            }
            __syncthreads();
        }
    }

    // Radix-16 stage on the matrix cores (16-bit storage).  The tile's
    // tpw * n/16 butterflies are flattened (slot-major, or slot-minor for
    // column-strided HBM access) and taken 32 at a time by a wave: lane l
    // holds points mfma_point(i, l / 32) of butterfly l % 32 of its group.
    // Compiled only into the MFMA instances (the 16-element one: 32 elements
    // would spill); the host enables it when blockDim is a multiple of 64,
    // the butterfly count a multiple of 32 and at most EPT / 8 groups per
    // wave.
    __device__ __forceinline__ void stage16_mfma(int s, bool src_global, bool dst_global,
                                                 bool col_map) {
        if constexpr (!St::native && MFMA) {
            constexpr int GM = EPT / 8;
            const StageDesc st = a.pd.st[s];
            const int nb = n_ / 16, ns = st.ns;
            const int lane = tid & 63, wv = tid >> 6, nw = (int)(blockDim.x >> 6);
            const int col = lane & 31, hh = lane >> 5;
            const int ngroups = (tpw_ * nb) >> 5;
            C2<T> v[GM][8];
            int gslot[GM], gj[GM];
#pragma unroll
            for (int q = 0; q < GM; ++q) {
                const int gi = wv + q * nw;
                gslot[q] = 0;
                gj[q] = 0;
#pragma unroll
                for (int i = 0; i < 8; ++i) v[q][i] = {(T)0, (T)0};
                if (gi < ngroups) {  // wave-uniform
                    const int f = gi * 32 + col;
                    const int slot = col_map ? f % tpw_ : f / nb;
                    const int j = col_map ? f / tpw_ : f % nb;
                    gslot[q] = slot;
                    gj[q] = j;
                    const int g = slot0 + slot;
                    if (src_global) {
                        const bool act = g < total_slots;
                        const int64_t base = act ? ibase(g) : 0;
                        const int64_t es = iin_.elem_stride;
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            C2<T> x = {(T)0, (T)0};
                            if (act) x = St::load(in, base + (int64_t)(j + mfma_in_point(i, hh) * nb) * es);
                            if (conj_in_) x.y = -x.y;
                            v[q][i] = x;
                        }
                    } else {
                        const C2<T> *ls = lds + slot * lds_stride_;
#pragma unroll
                        for (int i = 0; i < 8; ++i) v[q][i] = ls[lds_idx(j + mfma_in_point(i, hh) * nb)];
                    }
                }
            }
            if (!src_global) __syncthreads();
            const Dft16Operands<S> dm(lane);
#pragma unroll
            for (int q = 0; q < GM; ++q) {
                if (wv + q * nw < ngroups) {
                    if (ns > 1) {
                        const int k = gj[q] % ns;
                        const C2<T> *w = tw_ + st.tw_off + k * tw_group(16);
#pragma unroll
                        for (int i = 0; i < 8; ++i) {
                            const int r = mfma_in_point(i, hh);
                            if (r) v[q][i] = cmul(v[q][i], w[r - 1]);
                        }
                    }
                }
            }
            // every group, active or not (inactive ones hold zeros and are
            // never stored), two at a time: their MFMAs interleave
#pragma unroll
            for (int q = 0; q < GM; q += 2)
                dft16_mfma<S, 2>(*reinterpret_cast<C2<T>(*)[2][8]>(&v[q]), dm);
#pragma unroll
            for (int q = 0; q < GM; ++q) {
                if (wv + q * nw < ngroups) {
                    const int j = gj[q], slot = gslot[q];
                    const int qq = j / ns, k = j - qq * ns;
                    const int o = qq * ns * 16 + k;
                    const int g = slot0 + slot;
                    if (dst_global) {
                        if (g < total_slots) {
                            const int64_t base = obase(g);
                            const int64_t es = iout_.elem_stride;
#pragma unroll
                            for (int i = 0; i < 8; ++i) {
                                const int e = o + mfma_point(i, hh) * ns;
                                C2<T> y = v[q][i];
                                if (conj_out_) y.y = -y.y;
                                if (post_tw_) y = cmul(y, post_twiddle(g, e));
                                St::store(out, base + (int64_t)e * es, y);
                            }
                        }
                    } else {
                        C2<T> *ls = lds + slot * lds_stride_;
#pragma unroll
                        for (int i = 0; i < 8; ++i) ls[lds_idx(o + mfma_point(i, hh) * ns)] = v[q][i];
                    }
                }
            }
            if (!dst_global) __syncthreads();
        }
    }

    __device__ __forceinline__ void dispatch(int s, bool src_global, bool dst_global, bool col_map) {
        switch (a.pd.st[s].radix) {
        case 1: stage<1>(s, src_global, dst_global, col_map); break;
        case 2: stage<2>(s, src_global, dst_global, col_map); break;
        case 3: stage<3>(s, src_global, dst_global, col_map); break;
        case 4: stage<4>(s, src_global, dst_global, col_map); break;
        case 5: stage<5>(s, src_global, dst_global, col_map); break;
        case 6: stage<6>(s, src_global, dst_global, col_map); break;
        case 7: stage<7>(s, src_global, dst_global, col_map); break;
        case 8: stage<8>(s, src_global, dst_global, col_map); break;
        case 9: stage<9>(s, src_global, dst_global, col_map); break;
        case 10: stage<10>(s, src_global, dst_global, col_map); break;
        case 11: stage<11>(s, src_global, dst_global, col_map); break;
        case 12: stage<12>(s, src_global, dst_global, col_map); break;
        case 13: stage<13>(s, src_global, dst_global, col_map); break;
        case 15: stage<15>(s, src_global, dst_global, col_map); break;
        case 16: stage<16>(s, src_global, dst_global, col_map); break;
        case 25: stage<25>(s, src_global, dst_global, col_map); break;
        case 32:  // compiled 16384 plans; only in the 32-element instances
            if constexpr (EPT >= 32) stage<32>(s, src_global, dst_global, col_map);
            break;
        default: break;
        }
    }

    // Real inverse pre-process (pairs i, h - i) from HBM into LDS, writing
    // conj(.) so that the forward stages compute the inverse.
    __device__ __forceinline__ void real_pre() {
        if constexpr (!St::native) return;
        const int slot = tid / tpt_, lid = tid % tpt_;
        const int g = slot0 + slot;
        const bool active = (slot < tpw_) && (g < total_slots);
        const int h = n_;
        const bool mod = half_shift_ != 0;
        C2<T> *ls = lds + slot * lds_stride_;
        if (active) {
            const C2<T> *x = in + ibase(g);
            if (!mod && lid == 0) {
                const C2<T> v = x[0];
                ls[0] = {v.x + v.y, -(v.x - v.y)};
            }
            // kRealBatch pairs at a time, their loads issued before any use
            for (int i0 = (mod ? 0 : 1) + lid; i0 <= h / 2; i0 += kRealBatch * tpt_) {
                C2<T> va[kRealBatch], vb[kRealBatch], tw[kRealBatch];
#pragma unroll
                for (int k = 0; k < kRealBatch; ++k) {
                    const int i = i0 + k * tpt_;
                    if (i <= h / 2) {
                        const int ci = mod ? (h - 1 - i) : (h - i);
                        va[k] = x[i];
                        vb[k] = x[ci];
                        tw[k] = real_tw_[i];
                    }
                }
#pragma unroll
                for (int k = 0; k < kRealBatch; ++k) {
                    const int i = i0 + k * tpt_;
                    if (i > h / 2 || real_pair_skipped(i, h, mod)) continue;
                    const int ci = mod ? (h - 1 - i) : (h - i);
                    const C2<T> v = va[k], v2 = vb[k];
                    const C2<T> odd = {v.x + v2.x, v.y - v2.y};
                    const C2<T> erot = {v.x - v2.x, v.y + v2.y};
                    const C2<T> ev = cmulc(erot, tw[k]);
                    const C2<T> p = cadd(odd, ev);
                    const C2<T> m = csub(odd, ev);
                    ls[lds_idx(i)] = cconj(p);
                    ls[lds_idx(ci)] = m;  // conj(conj(odd - ev))
                }
            }
        }
        __syncthreads();
    }

    // Real forward post-process from the complex spectrum in LDS to HBM.
    __device__ __forceinline__ void real_post() {
        if constexpr (!St::native) return;
        const int slot = tid / tpt_, lid = tid % tpt_;
        const int g = slot0 + slot;
        const bool active = (slot < tpw_) && (g < total_slots);
        const int h = n_;
        const bool mod = half_shift_ != 0;
        const C2<T> *ls = lds + slot * lds_stride_;
        if (!active) return;
        C2<T> *y = out + obase(g);
        if (!mod && lid == 0) {
            const C2<T> z = ls[0];
            y[0] = {z.x + z.y, z.x - z.y};
        }
        const T half = (T)0.5;
        for (int i0 = (mod ? 0 : 1) + lid; i0 <= h / 2; i0 += kRealBatch * tpt_) {
            C2<T> za[kRealBatch], zb[kRealBatch], tw[kRealBatch];
#pragma unroll
            for (int k = 0; k < kRealBatch; ++k) {
                const int i = i0 + k * tpt_;
                if (i <= h / 2) {
                    const int ci = mod ? (h - 1 - i) : (h - i);
                    za[k] = ls[lds_idx(i)];
                    zb[k] = ls[lds_idx(ci)];
                    tw[k] = real_tw_[i];
                }
            }
#pragma unroll
            for (int k = 0; k < kRealBatch; ++k) {
                const int i = i0 + k * tpt_;
                if (i > h / 2 || real_pair_skipped(i, h, mod)) continue;
                const int ci = mod ? (h - 1 - i) : (h - i);
                const C2<T> zi = za[k], zc = zb[k];
                const C2<T> odd = {(zi.x + zc.x) * half, (zi.y - zc.y) * half};
                const C2<T> ev = {(zi.x - zc.x) * half, (zi.y + zc.y) * half};
                const C2<T> e = cmul(ev, tw[k]);
                y[i] = cadd(odd, e);
                y[ci] = cconj(csub(odd, e));
            }
        }
    }

    // LDS reads of the post-process done before the block reuses LDS.
    __device__ __forceinline__ void real_post_sync() {
        real_post();
        __syncthreads();
    }

    __device__ __forceinline__ void run() {
        const int last = nstages_ - 1;
        if constexpr (!St::native) {  // 16-bit storage: complex-to-complex only
            for (int s = 0; s <= last; ++s) {
                const bool col = (s == 0) ? (in_col_ != 0) : (s == last) ? (out_col_ != 0) : false;
                dispatch(s, s == 0, s == last, col);
            }
            return;
        }
        if (mode_ == kModeC2R) real_pre();
        for (int s = 0; s <= last; ++s) {
            const bool src_g = (s == 0) && (mode_ != kModeC2R);
            const bool dst_g = (s == last) && (mode_ != kModeR2C);
            const bool col = (s == 0) ? (in_col_ != 0) : (s == last) ? (out_col_ != 0) : false;
            dispatch(s, src_g, dst_g, col);
        }
        if (mode_ == kModeR2C) real_post_sync();
    }
};

// Forward real post-process in place over a spectrum z of length h (the
// output of a two-pass complex FFT): pair (i, ci) is read and written by one
// thread (signalsmith-fft.h:459-472).  Threads t0, t0 + nt, ... of the caller.
template <typename T>
__device__ __forceinline__ void real_post_inplace(C2<T> *z, const C2<T> *tmi, int h, bool mod,
                                                  int t0, int nt) {
    const T half = (T)0.5;
    for (int i = (mod ? 0 : 1) + t0; i <= h / 2; i += nt) {
        if (real_pair_skipped(i, h, mod)) continue;
        const int ci = mod ? (h - 1 - i) : (h - i);
        const C2<T> zi = z[i], zc = z[ci];
        const C2<T> odd = {(zi.x + zc.x) * half, (zi.y - zc.y) * half};
        const C2<T> ev = {(zi.x - zc.x) * half, (zi.y + zc.y) * half};
        const C2<T> e = cmul(ev, tmi[i]);
        z[i] = cadd(odd, e);
        z[ci] = cconj(csub(odd, e));
    }
    if (!mod && t0 == 0) {
        const C2<T> d = z[0];
        z[0] = {d.x + d.y, d.x - d.y};
    }
}

}  // namespace sfft
