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

// Threads per persistent workgroup: 1024 for float; 512 for double, whose
// butterflies need more than the 128 registers a 1024-thread block allows.
template <typename T>
constexpr int persistent_block() {
    return sizeof(T) == 8 ? 512 : 1024;
}
// WorkUnit.npasses bit: forward real two-pass unit, post-process after pass 2
constexpr int32_t kUnitRealPost = 0x100;

// Launches the persistent kernel: `grid` workgroups pull units from a
// device counter until the list is exhausted; the last workgroup to finish
// resets the counters, so consecutive launches (and graph replays) need no
// memset.  counters: 2 x uint32, zero before the first launch.
template <typename T>
hipError_t launch_persistent(const WorkUnit *units, int nunits, const PassArgs<T> *table,
                             const int64_t *gather_in, const int64_t *gather_out,
                             unsigned *counters, const void *in_arena, void *out_arena,
                             unsigned grid, size_t lds_bytes, hipStream_t stream);

template <typename T>
int persistent_blocks_per_cu(size_t lds_bytes);

}  // namespace sfft
