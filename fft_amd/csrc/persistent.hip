// persistent.hip -- one launch over an uneven batch of transform sizes.
#include "persistent.h"

namespace sfft {

template <typename T, int EPT>
__global__ void __launch_bounds__(persistent_block<T>())
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

template <typename T>
static constexpr int persistent_ept() {
    return sizeof(T) == 4 ? 32 : 16;
}

template <typename T>
hipError_t launch_persistent(const WorkUnit *units, int nunits, const PassArgs<T> *table,
                             const int64_t *gather_in, const int64_t *gather_out,
                             unsigned *counters, const void *in_arena, void *out_arena,
                             unsigned grid, size_t lds_bytes, hipStream_t stream) {
    constexpr int E = persistent_ept<T>();
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_persistent<T, E>) != hipSuccess) return lds_err;
    hipLaunchKernelGGL((k_persistent<T, E>), dim3(grid), dim3(persistent_block<T>()), lds_bytes,
                       stream, units, nunits, table, gather_in, gather_out, counters,
                       (const char *)in_arena,
                       (char *)out_arena);
    return hipGetLastError();
}

template <typename T>
int persistent_blocks_per_cu(size_t lds_bytes) {
    constexpr int E = persistent_ept<T>();
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, (const void *)k_persistent<T, E>,
                                                     persistent_block<T>(), lds_bytes) != hipSuccess)
        return 1;
    return blocks > 0 ? blocks : 1;
}

template hipError_t launch_persistent<float>(const WorkUnit *, int, const PassArgs<float> *,
                                             const int64_t *, const int64_t *, unsigned *,
                                             const void *, void *, unsigned, size_t, hipStream_t);
template hipError_t launch_persistent<double>(const WorkUnit *, int, const PassArgs<double> *,
                                              const int64_t *, const int64_t *, unsigned *,
                                              const void *, void *, unsigned, size_t,
                                              hipStream_t);
template int persistent_blocks_per_cu<float>(size_t);
template int persistent_blocks_per_cu<double>(size_t);

}  // namespace sfft
