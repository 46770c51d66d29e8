// fused32.hip -- FP32 instantiation of the fused transport kernel (built
// with FMA contraction on: the FP32 form is compared at 1e-5, not bitwise).
#include "fused.h"
#include "fused_kernel.h"

namespace fv3 {

int launch_fused_f32(const FusedArgs<float> &a, int nblk, hipStream_t s) {
  dim3 g((unsigned)nblk, (unsigned)a.npz, 1), b(FT_LANES * FUSED32_WAVES);
  hipLaunchKernelGGL((k_fused_tp<float, FUSED32_WAVES>), g, b, 0, s, a);
  FV3_HIP(hipGetLastError());
  return 0;
}

}  // namespace fv3
