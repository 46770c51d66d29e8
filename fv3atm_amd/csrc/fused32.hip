// fused32.hip -- FP32 instantiations of the fused transport kernel (built
// with FMA contraction on: the FP32 form is compared at 1e-5, not bitwise).
#include "fused.h"
#include "fused_kernel.h"

namespace fv3 {

template <int W, int NT, int HO>
static void go2(const FusedArgs<float> &a, int nblk, hipStream_t s) {
  const dim3 b(FT_LANES * W);
  const int nin = a.s_hi >= a.s_lo && a.v_hi >= a.v_lo ? (a.s_hi - a.s_lo + 1) * (a.v_hi - a.v_lo + 1) : 0;
  if (nin > 0)
    hipLaunchKernelGGL((k_fused_tp<float, W, NT, HO, false>), dim3((unsigned)nin, (unsigned)a.npz, (unsigned)a.ntiles), b, 0, s, a);
  if (nin < nblk)
    hipLaunchKernelGGL((k_fused_tp<float, W, NT, HO, true>), dim3((unsigned)nblk, (unsigned)a.npz, (unsigned)a.ntiles), b, 0, s, a);
}

template <int W, int NT>
static int go(const FusedArgs<float> &a, int nblk, hipStream_t s) {
  if (a.hord == 8)  // the production order, specialised at compile time
    go2<W, NT, 8>(a, nblk, s);
  else
    go2<W, NT, 0>(a, nblk, s);
  FV3_HIP(hipGetLastError());
  return 0;
}

int launch_fused_f32(const FusedArgs<float> &a, int nblk, int cfg, hipStream_t s) {
  switch (cfg) {
    case 0: return go<4, 1>(a, nblk, s);
    case 1: return go<8, 1>(a, nblk, s);
    case 2: return go<4, 2>(a, nblk, s);
    case 3: return go<2, 2>(a, nblk, s);
    case 4: return go<6, 1>(a, nblk, s);
    case 5: return go<1, 2>(a, nblk, s);
    case 6: return go<1, 1>(a, nblk, s);
    case 7: return go<2, 1>(a, nblk, s);
    default: return set_error("unknown fused FP32 launch configuration %d", cfg);
  }
}

}  // namespace fv3
