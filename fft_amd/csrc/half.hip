// half.hip -- pass kernel instances for 16-bit HBM storage (IEEE half,
// bfloat16) with float arithmetic and radix-16 stages on MFMA.
#include "kernels.h"

namespace sfft {

// The float-butterfly instances keep 4 waves per SIMD (at most 128
// registers); the MFMA instance needs more for its split operands.
template <int EPT, typename S, bool MFMA>
__global__ void __launch_bounds__(kMaxBlock16, MFMA ? 1 : 4) k_pass16(PassArgs<float> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<float> *lds = reinterpret_cast<C2<float> *>(smem_raw);
    Block<float, EPT, S, MFMA> blk(a, a.in, a.out, a.total_slots, lds, (int)threadIdx.x,
                             (int)xcd_tile(blockIdx.x, gridDim.x) * a.pd.tpw);
    blk.run();
}

template <int EPT, typename S, bool MFMA>
static hipError_t launch16(const PassArgs<float> &a, unsigned grid, unsigned block,
                           size_t lds_bytes, hipStream_t stream) {
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_pass16<EPT, S, MFMA>) != hipSuccess) return lds_err;
    hipLaunchKernelGGL((k_pass16<EPT, S, MFMA>), dim3(grid), dim3(block), lds_bytes, stream, a);
    return hipGetLastError();
}

hipError_t launch_pass_16(const PassArgs<float> &a, int storage, int ept, unsigned grid,
                          unsigned block, size_t lds_bytes, hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    // MFMA radix-16 stages: the 16-element instance only
    const bool mfma = a.pd.mfma != 0;
    if (mfma && ept != kEptSmall) return hipErrorInvalidValue;
    const bool small = ept <= kEptSmall;
    if (storage == kStorageF16) {
        if (mfma) return launch16<kEptSmall, StoreF16, true>(a, grid, block, lds_bytes, stream);
        return small ? launch16<kEptSmall, StoreF16, false>(a, grid, block, lds_bytes, stream)
                     : launch16<kEptLarge, StoreF16, false>(a, grid, block, lds_bytes, stream);
    }
    if (storage == kStorageBF16) {
        if (mfma) return launch16<kEptSmall, StoreBF16, true>(a, grid, block, lds_bytes, stream);
        return small ? launch16<kEptSmall, StoreBF16, false>(a, grid, block, lds_bytes, stream)
                     : launch16<kEptLarge, StoreBF16, false>(a, grid, block, lds_bytes, stream);
    }
    return hipErrorInvalidValue;
}

}  // namespace sfft
