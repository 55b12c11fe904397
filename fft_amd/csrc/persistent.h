// persistent.h -- work units of the persistent uneven-batch kernel.
#pragma once

#include <hip/hip_runtime.h>

#include "stockham.h"

namespace sfft {

// A unit: `count` transforms of one plan, contiguous in the arenas.
// Single-pass units gather up to one round of same-plan transforms from
// arbitrary arena positions (gather >= 0: element offsets at gather_in /
// gather_out[gather + s]); two-pass units hold one transform at in_off/out_off.
struct WorkUnit {
    int32_t args;     // first PassArgs entry of the unit's plan and direction
    int32_t npasses;  // 1 (LDS resident) or 2 (column pass, then in-place pass)
    int32_t count;    // transforms in this unit
    int32_t gather;   // first entry in the gather lists, or -1
    int64_t in_off;   // byte offsets into the input / output arenas
    int64_t out_off;
};

// Kernel instances of the persistent launch ("classes"): threads per block,
// elements per thread, and blocks per CU (each block may use that share of
// the CU's LDS).  Every float class holds 16 waves per CU at up to 128
// registers each, and a unit runs in the first class that holds it: class 0
// runs two 512-thread blocks per CU, so one block's loads and stores overlap
// the other's stages.  (Four 256-thread or eight 128-thread blocks per CU
// measured slower on the uneven batch, 1.98 -> 2.27 ms: the same bytes in
// flight per CU, in more and shorter rounds.)  FP64 (whose butterflies need
// more than 128 registers) has one class.
struct PersistentClass {
    int threads;
    int ept;
    int blocks_per_cu;
};
constexpr int kPersistentClasses = 3;
template <typename T>
inline int persistent_classes() {
    return sizeof(T) == 8 ? 1 : kPersistentClasses;
}
template <typename T>
inline PersistentClass persistent_class(int c) {
    if (sizeof(T) == 8) return {512, 16, 1};
    constexpr PersistentClass k[kPersistentClasses] = {{512, 16, 2}, {1024, 16, 1}, {1024, 32, 1}};
    return k[c];
}
// WorkUnit.npasses bit: forward real two-pass unit, post-process after pass 2
constexpr int32_t kUnitRealPost = 0x100;

// Launches the persistent kernel of class cls: `grid` workgroups pull units
// from a device counter until the list is exhausted; the last workgroup to
// finish resets the counters, so consecutive launches (and graph replays)
// need no memset.  counters: 2 x uint32, zero before the first launch.
template <typename T>
hipError_t launch_persistent(int cls, const WorkUnit *units, int nunits, const PassArgs<T> *table,
                             const int64_t *gather_in, const int64_t *gather_out,
                             unsigned *counters, const void *in_arena, void *out_arena,
                             unsigned grid, size_t lds_bytes, hipStream_t stream);

template <typename T>
int persistent_blocks_per_cu(int cls, size_t lds_bytes);

}  // namespace sfft
