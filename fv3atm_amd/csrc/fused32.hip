// fused32.hip -- FP32 instantiations of the fused transport kernel (built
// with FMA contraction on: the FP32 form is compared at 1e-5, not bitwise).
#include "fused.h"
#include "fused_kernel.h"

namespace fv3 {

// W waves per workgroup x NT tracers per wave for the lean kind, WE x NTE
// for the two edge kinds (separate launches over disjoint blocks); OCC_E:
// occupancy asked of the edge kinds' register allocation (the lean kind with
// one tracer per wave fits three waves per SIMD without spilling)
template <int W, int NT, int HO, int OCC_E, int WE, int NTE, int OCC_L0>
static void go2(const FusedArgs<float> &a, const KindStreams &s) {
  constexpr int OCC_L = OCC_L0 > 0 ? OCC_L0 : (NT > 1 ? 1 : (W >= 4 ? 12 / W : 3));
  int n[7];
  fused_grid_split(a, n);
  auto grid = [&](int c) { return dim3((unsigned)c, (unsigned)a.npz, (unsigned)a.ntiles); };
  const dim3 bl(FT_LANES * W), be(FT_LANES * WE);
  if (n[0] > 0) hipLaunchKernelGGL((k_fused_tp<float, W, NT, HO, 0, OCC_L>), grid(n[0]), bl, 0, s.k[0], a);
  if (n[1] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 1, OCC_E>), grid(n[1]), be, 0, s.k[1], a);
  if (n[2] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 2, OCC_E>), grid(n[2]), be, 0, s.k[2], a);
  if (n[3] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 3, OCC_E>), grid(n[3]), be, 0, s.k[2], a);
  if (n[4] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 4, OCC_E>), grid(n[4]), be, 0, s.k[3], a);
  if (n[5] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 5, OCC_E>), grid(n[5]), be, 0, s.k[2], a);
  if (n[6] > 0) hipLaunchKernelGGL((k_fused_tp<float, WE, NTE, HO, 6, OCC_E>), grid(n[6]), be, 0, s.k[3], a);
}

template <int W, int NT, int OCC_E = 1, int WE = W, int NTE = NT, int OCC_L = 0>
static int go(const FusedArgs<float> &a, const KindStreams &s) {
  if (a.hord == 8)  // the production order, specialised at compile time
    go2<W, NT, 8, OCC_E, WE, NTE, OCC_L>(a, s);
  else
    go2<W, NT, 0, OCC_E, WE, NTE, OCC_L>(a, s);
  FV3_HIP(hipGetLastError());
  return 0;
}

// W x NT of the lean kind of each case below
static const int kRound[] = {8, 8, 4, 4, 6, 2, 1, 2, 4, 8, 4, 8, 16, 12, 8};
static_assert(sizeof(kRound) / sizeof(kRound[0]) == FUSED32_NCFG, "one entry per launch configuration");

int fused32_round(int cfg) { return cfg >= 0 && cfg < FUSED32_NCFG ? kRound[cfg] : 8; }

int launch_fused_f32(const FusedArgs<float> &a, int cfg, const KindStreams &s) {
  switch (cfg) {
    case 0: return go<4, 2>(a, s);
    case 1: return go<8, 1>(a, s);
    case 2: return go<4, 1>(a, s);
    case 3: return go<2, 2>(a, s);
    case 4: return go<6, 1>(a, s);
    case 5: return go<1, 2>(a, s);
    case 6: return go<1, 1>(a, s);
    case 7: return go<2, 1>(a, s);
    case 8: return go<4, 1, 3>(a, s);
    case 9: return go<4, 2, 3, 4, 1>(a, s);
    case 10: return go<4, 1, 3, 4, 1, 4>(a, s);
    case 11: return go<4, 2, 3, 4, 1, 3>(a, s);
    case 12: return go<8, 2, 3, 4, 1>(a, s);
    case 13: return go<6, 2, 3, 4, 1>(a, s);
    case 14: return go<8, 1, 3, 4, 1>(a, s);
    default: return set_error("unknown fused FP32 launch configuration %d", cfg);
  }
}

}  // namespace fv3
