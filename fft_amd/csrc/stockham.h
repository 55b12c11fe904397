// stockham.h -- LDS-resident Stockham block FFT for gfx950.
//
// One workgroup holds `tpw` slots (sub-transforms) of length n in LDS.  The
// first stage reads HBM directly (coalesced: consecutive lanes take
// consecutive elements, or consecutive slots for column-strided input), every
// stage boundary is one LDS exchange, and the last stage writes HBM directly.
// The inverse transform is conj(FFT(conj(x))): conj on load and on store.
#pragma once

#include "dft.h"
#include "plan_desc.h"

namespace sfft {

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv &f) {
    if (f.d <= 1) return n;
    return (__umulhi(n, f.mul) + n) >> f.shift;
}

__device__ __forceinline__ int lds_idx(int e) { return e + (e >> kLdsPadShift); }

// Real-FFT fusion modes.
enum : int32_t {
    kModeC2C = 0,
    kModeR2C = 1,  // forward real: post-process pairs (i, h-i) after the last stage
    kModeC2R = 2,  // inverse real: pre-process pairs before the first stage
};

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
    const C2<double> *ptw_fine;    // W_N^{m}, m < 2^shift
    const C2<double> *ptw_coarse;  // W_N^{m 2^shift}
    int32_t in_from_out;   // second pass of a two-pass plan: reads the output buffer
    int32_t pad_;
};

// The pass descriptor `a` may live in kernel arguments or in a device table
// (persistent launches); in, out and total_slots are per launch / per item.
// Optional gather lists (persistent launches) give each slot its own input /
// output offset in elements instead of the strided address map.
template <typename T, int EPT>
struct Block {
    const PassArgs<T> &a;
    const C2<T> *in;
    C2<T> *out;
    int total_slots;
    C2<T> *lds;
    int tid;
    int slot0;
    const int64_t *gather_in = nullptr;
    const int64_t *gather_out = nullptr;

    __device__ Block(const PassArgs<T> &args, const C2<T> *in_, C2<T> *out_, int total, C2<T> *l,
                     int t, int s0)
        : a(args), in(in_), out(out_), total_slots(total), lds(l), tid(t), slot0(s0) {}

    __device__ __forceinline__ int64_t ibase(int g) const {
        return gather_in ? gather_in[g] : gbase(a.iin, g);
    }
    __device__ __forceinline__ int64_t obase(int g) const {
        return gather_out ? gather_out[g] : gbase(a.iout, g);
    }

    __device__ __forceinline__ int64_t gbase(const IoDesc &io, int g) const {
        if (io.per_batch.d <= 1) return (int64_t)g * io.batch_stride;
        uint32_t b = fdiv((uint32_t)g, io.per_batch);
        uint32_t r = (uint32_t)g - b * io.per_batch.d;
        return (int64_t)b * io.batch_stride + (int64_t)r * io.slot_stride;
    }

    __device__ __forceinline__ C2<T> post_twiddle(int g, int e) const {
        // W_N^{c e}, c = column index of slot g, via a two-level FP64 table.
        const uint32_t c = (uint32_t)g - fdiv((uint32_t)g, a.iout.per_batch) * a.iout.per_batch.d;
        // c < N2 and e < N1, so c e < N: no reduction needed.
        const uint32_t m = c * (uint32_t)e;
        const C2<double> f = a.ptw_fine[m & ((1u << a.post_tw_shift) - 1u)];
        const C2<double> h = a.ptw_coarse[m >> a.post_tw_shift];
        return {(T)fmadd(f.x, h.x, -(f.y * h.y)), (T)fmadd(f.x, h.y, f.y * h.x)};
    }

    // One Stockham stage of radix R.
    template <int R>
    __device__ __forceinline__ void stage(int s, bool src_global, bool dst_global, bool col_map) {
        const PassDesc &pd = a.pd;
        const StageDesc st = pd.st[s];
        constexpr int B = (EPT / R) > 0 ? (EPT / R) : 1;
        const int n = pd.n;
        const int nb = n / R;
        int slot, lid;
        if (col_map) {
            slot = tid % pd.tpw;
            lid = tid / pd.tpw;
        } else {
            slot = tid / pd.tpt;
            lid = tid % pd.tpt;
        }
        const int g = slot0 + slot;
        const bool active = (slot < pd.tpw) && (lid < pd.tpt) && (g < total_slots);
        C2<T> *ls = lds + slot * pd.lds_stride;

        C2<T> v[B][R];
        // ---- load
        if (src_global) {
            const int64_t base = active ? ibase(g) : 0;
            const int64_t es = a.iin.elem_stride;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * pd.tpt;
                if (active && j < nb) {
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int e = j + r * nb;
                        C2<T> x = in[base + (int64_t)e * es];
                        if (a.conj_in) x.y = -x.y;
                        if (a.mode == kModeR2C && a.half_shift) x = cmul(x, a.rot[e]);
                        v[b][r] = x;
                    }
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * pd.tpt;
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
            const int j = lid + b * pd.tpt;
            if (active && j < nb) {
                if (ns > 1) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const C2<T> *w = a.tw + st.tw_off + k * tw_group(R);
#pragma unroll
                    for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], w[r - 1]);
                }
                dft<R, T>(v[b]);
            }
        }
        // ---- store
        if (dst_global) {
            const int64_t base = active ? obase(g) : 0;
            const int64_t es = a.iout.elem_stride;
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * pd.tpt;
                if (active && j < nb) {
                    const uint32_t q = fdiv((uint32_t)j, st.div);
                    const int k = j - (int)q * ns;
                    const int o = (int)q * ns * R + k;
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const int e = o + r * ns;
                        C2<T> y = v[b][r];
                        if (a.conj_out) y.y = -y.y;
                        if (a.mode == kModeC2R && a.half_shift) y = cmulc(y, a.rot[e]);
                        if (a.post_tw) y = cmul(y, post_twiddle(g, e));
                        out[base + (int64_t)e * es] = y;
                    }
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const int j = lid + b * pd.tpt;
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
        default: break;
        }
    }

    // Real inverse pre-process (pairs i, h - i) from HBM into LDS, writing
    // conj(.) so that the forward stages compute the inverse.
    __device__ __forceinline__ void real_pre() {
        const PassDesc &pd = a.pd;
        const int slot = tid / pd.tpt, lid = tid % pd.tpt;
        const int g = slot0 + slot;
        const bool active = (slot < pd.tpw) && (g < total_slots);
        const int h = pd.n;
        const bool mod = a.half_shift != 0;
        C2<T> *ls = lds + slot * pd.lds_stride;
        if (active) {
            const C2<T> *x = in + ibase(g);
            if (!mod && lid == 0) {
                const C2<T> v = x[0];
                ls[0] = {v.x + v.y, -(v.x - v.y)};
            }
            for (int i = (mod ? 0 : 1) + lid; i <= h / 2; i += pd.tpt) {
                const int ci = mod ? (h - 1 - i) : (h - i);
                const C2<T> v = x[i], v2 = x[ci];
                const C2<T> odd = {v.x + v2.x, v.y - v2.y};
                const C2<T> erot = {v.x - v2.x, v.y + v2.y};
                const C2<T> ev = cmulc(erot, a.real_tw[i]);
                const C2<T> p = cadd(odd, ev);
                const C2<T> m = csub(odd, ev);
                ls[lds_idx(i)] = cconj(p);
                ls[lds_idx(ci)] = m;  // conj(conj(odd - ev))
            }
        }
        __syncthreads();
    }

    // Real forward post-process from the complex spectrum in LDS to HBM.
    __device__ __forceinline__ void real_post() {
        const PassDesc &pd = a.pd;
        const int slot = tid / pd.tpt, lid = tid % pd.tpt;
        const int g = slot0 + slot;
        const bool active = (slot < pd.tpw) && (g < total_slots);
        const int h = pd.n;
        const bool mod = a.half_shift != 0;
        const C2<T> *ls = lds + slot * pd.lds_stride;
        if (!active) return;
        C2<T> *y = out + obase(g);
        if (!mod && lid == 0) {
            const C2<T> z = ls[0];
            y[0] = {z.x + z.y, z.x - z.y};
        }
        const T half = (T)0.5;
        for (int i = (mod ? 0 : 1) + lid; i <= h / 2; i += pd.tpt) {
            const int ci = mod ? (h - 1 - i) : (h - i);
            const C2<T> zi = ls[lds_idx(i)], zc = ls[lds_idx(ci)];
            const C2<T> odd = {(zi.x + zc.x) * half, (zi.y - zc.y) * half};
            const C2<T> ev = {(zi.x - zc.x) * half, (zi.y + zc.y) * half};
            const C2<T> e = cmul(ev, a.real_tw[i]);
            y[i] = cadd(odd, e);
            y[ci] = cconj(csub(odd, e));
        }
    }

    // LDS reads of the post-process done before the block reuses LDS.
    __device__ __forceinline__ void real_post_sync() {
        real_post();
        __syncthreads();
    }

    __device__ __forceinline__ void run() {
        const PassDesc &pd = a.pd;
        const int last = pd.nstages - 1;
        if (a.mode == kModeC2R) real_pre();
        for (int s = 0; s <= last; ++s) {
            const bool src_g = (s == 0) && (a.mode != kModeC2R);
            const bool dst_g = (s == last) && (a.mode != kModeR2C);
            const bool col = (s == 0) ? (pd.in_col != 0) : (s == last) ? (pd.out_col != 0) : false;
            dispatch(s, src_g, dst_g, col);
        }
        if (a.mode == kModeR2C) real_post_sync();
    }
};

}  // namespace sfft
