// persistent.hip -- one launch over an uneven batch of transform sizes.
#include "persistent.h"

namespace sfft {

// MINW: waves per SIMD the register allocation must allow (4: 128 registers).
template <typename T, int EPT, int THREADS, int MINW>
__global__ void __launch_bounds__(THREADS, MINW)
    k_persistent(const WorkUnit *__restrict__ units, int nunits,
                 const PassArgs<T> *__restrict__ table, const int64_t *__restrict__ gather_in,
                 const int64_t *__restrict__ gather_out, unsigned *counters,
                 const char *__restrict__ in_arena, char *__restrict__ out_arena) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_unit;
    C2<T> *lds = reinterpret_cast<C2<T> *>(smem_raw);
    const int tid = (int)threadIdx.x;
    for (;;) {
        if (tid == 0) s_unit = (int)atomicAdd(&counters[0], 1u);
        __syncthreads();
        const int u = __builtin_amdgcn_readfirstlane(s_unit);
        __syncthreads();  // s_unit consumed before thread 0 may overwrite it
        if (u >= nunits) break;
        const WorkUnit wu = units[u];
        const bool gathered = wu.gather >= 0;
        const C2<T> *in = reinterpret_cast<const C2<T> *>(in_arena + (gathered ? 0 : wu.in_off));
        C2<T> *out = reinterpret_cast<C2<T> *>(out_arena + (gathered ? 0 : wu.out_off));
        const int npasses = wu.npasses & 0xff;
        for (int p = 0; p < npasses; ++p) {
            const PassArgs<T> &a = table[wu.args + p];
            const int spt = (int)a.iin.per_batch.d > 1 ? (int)a.iin.per_batch.d : 1;
            const int total = wu.count * spt;
            const C2<T> *src = a.in_from_out ? out : in;
            for (int s0 = 0; s0 < total; s0 += a.pd.tpw) {
                Block<T, EPT> blk(a, src, out, total, lds, tid, s0);
                if (gathered) {
                    blk.gather_in = gather_in + wu.gather;
                    blk.gather_out = gather_out + wu.gather;
                }
                blk.run();
            }
            if (p + 1 < npasses || (wu.npasses & kUnitRealPost)) {
                // The next pass is read by this same workgroup (same CU), and no
                // other workgroup touches this unit's output: a workgroup-scope
                // fence (stores complete) suffices.  A device-scope fence here
                // would write back the whole L2 of the XCD for every unit.
                __threadfence_block();
                __syncthreads();
            }
        }
        if (wu.npasses & kUnitRealPost) {
            const PassArgs<T> &a = table[wu.args + npasses - 1];
            real_post_inplace(out, a.real_tw, a.real_h, a.half_shift != 0, tid,
                              (int)blockDim.x);
            __syncthreads();
        }
    }
    // Every workgroup has made its final (failing) grab; the last one to
    // arrive resets both counters for the next launch.
    if (tid == 0) {
        __threadfence();
        const unsigned done = atomicAdd(&counters[1], 1u);
        if (done == gridDim.x - 1) {
            atomicExch(&counters[0], 0u);
            atomicExch(&counters[1], 0u);
        }
    }
}

template <typename T, int E, int TH, int MW>
static hipError_t launch_one(const WorkUnit *units, int nunits, const PassArgs<T> *table,
                             const int64_t *gather_in, const int64_t *gather_out,
                             unsigned *counters, const void *in_arena, void *out_arena,
                             unsigned grid, size_t lds_bytes, hipStream_t stream) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_persistent<T, E, TH, MW>) != hipSuccess)
        return lds_err;
    hipLaunchKernelGGL((k_persistent<T, E, TH, MW>), dim3(grid), dim3(TH), lds_bytes, stream,
                       units, nunits, table, gather_in, gather_out, counters,
                       (const char *)in_arena, (char *)out_arena);
    return hipGetLastError();
}

template <typename T, int E, int TH, int MW>
static int occupancy(size_t lds_bytes) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    int blocks = 0;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_persistent<T, E, TH, MW>) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)k_persistent<T, E, TH, MW>,
                                                     TH, lds_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return 1;
    }
    return blocks > 0 ? blocks : 1;
}

// the instance of each float class (persistent_class): elements per thread,
// threads, waves per SIMD
template <>
hipError_t launch_persistent<float>(int cls, const WorkUnit *units, int nunits,
                                    const PassArgs<float> *table, const int64_t *gather_in,
                                    const int64_t *gather_out, unsigned *counters,
                                    const void *in_arena, void *out_arena, unsigned grid,
                                    size_t lds_bytes, hipStream_t stream) {
    switch (cls) {
#define SFFT_CASE(C, E, TH)                                                                     \
    case C:                                                                                   \
        return launch_one<float, E, TH, 4>(units, nunits, table, gather_in, gather_out, counters, \
                                           in_arena, out_arena, grid, lds_bytes, stream);
        SFFT_CASE(0, 16, 512)
        SFFT_CASE(1, 16, 1024)
        SFFT_CASE(2, 32, 1024)
#undef SFFT_CASE
    default:
        return hipErrorInvalidValue;
    }
}

template <>
hipError_t launch_persistent<double>(int cls, const WorkUnit *units, int nunits,
                                     const PassArgs<double> *table, const int64_t *gather_in,
                                     const int64_t *gather_out, unsigned *counters,
                                     const void *in_arena, void *out_arena, unsigned grid,
                                     size_t lds_bytes, hipStream_t stream) {
    if (cls != 0) return hipErrorInvalidValue;
    return launch_one<double, 16, 512, 2>(units, nunits, table, gather_in, gather_out, counters,
                                          in_arena, out_arena, grid, lds_bytes, stream);
}

template <>
int persistent_blocks_per_cu<float>(int cls, size_t lds_bytes) {
    switch (cls) {
    case 0: return occupancy<float, 16, 512, 4>(lds_bytes);
    case 1: return occupancy<float, 16, 1024, 4>(lds_bytes);
    default: return occupancy<float, 32, 1024, 4>(lds_bytes);
    }
}
template <>
int persistent_blocks_per_cu<double>(int, size_t lds_bytes) {
    return occupancy<double, 16, 512, 2>(lds_bytes);
}

}  // namespace sfft
