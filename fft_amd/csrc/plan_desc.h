// plan_desc.h -- descriptors shared by the host planner and the kernels.
//
// A transform of length N runs as one or two "passes".  A pass is a batch of
// independent sub-transforms ("slots") of length n, each held by `tpt`
// threads of a workgroup that holds `tpw` slots in LDS.  Inside a pass the
// Stockham autosort stages run radix by radix:
//   stage s, radix R, span Ns (product of the earlier radices):
//     butterfly j in [0, n/R): k = j mod Ns,
//       reads  x[j + r n/R]                         r = 0..R-1
//       twiddles W_{Ns R}^{r k}
//       writes y[(j / Ns) Ns R + k + r Ns]
// so input and output are in natural order and no permutation pass exists.
#pragma once

#include <stdint.h>

namespace sfft {

constexpr int kMaxStages = 14;
constexpr int kLdsPadShift = 4;  // one pad element per 16 (breaks power-of-two strides)
constexpr int kLdsReserve = 256; // bytes of LDS kept for static __shared__ variables

// Unsigned division by an invariant divisor: q = (umulhi(n, mul) + n) >> shift.
// Exact for n < 2^31 (Granlund-Montgomery round-up method).
struct FastDiv {
    uint32_t d, mul, shift, pad;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f{d, 0u, 0u, 0u};
    if (d <= 1) return f;
    uint32_t l = 0;
    while ((1ull << l) < d) ++l;
    f.shift = l;
    f.mul = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
    return f;
}

// Twiddles of one butterfly (r = 1..R-1) are stored densely, R-1 entries per
// group.  (Groups padded to an even length for 16-byte pair loads were
// measured slower: the larger table costs more in cache than the wider
// loads save.)
constexpr int tw_group(int radix) { return radix - 1; }

struct StageDesc {
    int32_t radix;
    int32_t ns;      // span before this stage
    int32_t tw_off;  // offset of this stage's twiddles (ns groups of tw_group(radix))
    int32_t pad;
    FastDiv div;     // division by ns
};

struct PassDesc {
    int32_t n;           // sub-transform length
    int32_t nstages;
    int32_t tpt;         // threads per slot
    int32_t tpw;         // slots per workgroup
    int32_t lds_stride;  // padded LDS elements per slot
    int32_t in_col;      // stage 0 mapping: lanes along slots (strided elements)
    int32_t out_col;     // last stage mapping
    int32_t mfma;        // 16-bit storage: radix-16 stages as MFMA contractions
    StageDesc st[kMaxStages];
};

// Global address of element e of slot g:
//   (g / per_batch) * batch_stride + (g % per_batch) * slot_stride + e * elem_stride
struct IoDesc {
    int64_t batch_stride;
    int64_t slot_stride;
    int64_t elem_stride;
    FastDiv per_batch;
};

inline int32_t lds_padded(int32_t n) { return n + (n >> kLdsPadShift) + 1; }

}  // namespace sfft
