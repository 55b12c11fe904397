// f32x.hip -- pass kernel for FP32 HBM storage with FP64 twiddles, LDS and
// arithmetic: the data is rounded to FP32 only where it is stored (the
// output, and the intermediate of two-pass plans).
#include "kernels.h"

namespace sfft {

__global__ void __launch_bounds__(kMaxBlock<double>()) k_pass_f32x(PassArgs<double> a) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2<double> *lds = reinterpret_cast<C2<double> *>(smem_raw);
    Block<double, kEptSmall, StoreF32> blk(a, a.in, a.out, a.total_slots, lds, (int)threadIdx.x,
                                           (int)xcd_tile(blockIdx.x, gridDim.x) * a.pd.tpw);
    blk.run();
}

hipError_t launch_pass_f32x(const PassArgs<double> &a, unsigned grid, unsigned block,
                            size_t lds_bytes, hipStream_t stream) {
    if (grid == 0) return hipSuccess;
    static std::once_flag lds_once;
    static hipError_t lds_err = hipSuccess;
    if (allow_max_lds(lds_once, lds_err, (const void *)k_pass_f32x) != hipSuccess) return lds_err;
    hipLaunchKernelGGL(k_pass_f32x, dim3(grid), dim3(block), lds_bytes, stream, a);
    return hipGetLastError();
}

}  // namespace sfft
