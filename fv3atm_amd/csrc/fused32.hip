// fused32.hip -- FP32 instantiations of the fused transport kernel (built
// with FMA contraction on: the FP32 form is compared at 1e-5, not bitwise).
#include "fused.h"
#include "fused_kernel.h"

namespace fv3 {

template <int W, int NT>
static int go(const FusedArgs<float> &a, int nblk, hipStream_t s) {
  dim3 g((unsigned)nblk, (unsigned)a.npz, 1), b(FT_LANES * W);
  hipLaunchKernelGGL((k_fused_tp<float, W, NT>), g, b, 0, s, a);
  FV3_HIP(hipGetLastError());
  return 0;
}

int launch_fused_f32(const FusedArgs<float> &a, int nblk, int cfg, hipStream_t s) {
  switch (cfg) {
    case 0: return go<5, 2>(a, nblk, s);
    case 1: return go<8, 1>(a, nblk, s);
    case 2: return go<4, 2>(a, nblk, s);
    case 3: return go<3, 2>(a, nblk, s);
    case 4: return go<2, 3>(a, nblk, s);
    case 5: return go<1, 2>(a, nblk, s);
    default: return set_error("unknown fused FP32 launch configuration %d", cfg);
  }
}

}  // namespace fv3
