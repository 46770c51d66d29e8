// fused32.hip -- FP32 instantiations of the fused transport kernel (built
// with FMA contraction on: the FP32 form is compared at 1e-5, not bitwise).
//
// The launch configurations are instantiated over FUSED32_NPART objects so
// that they compile in parallel: the build compiles this file once per part
// with -DFV3_F32_PART=p, and part p instantiates the configurations cfg with
// cfg % FUSED32_NPART == p.  Part 0 also holds the dispatcher.
#include "fused.h"
#include "fused_kernel.h"

#ifndef FV3_F32_PART
#error "fused32.hip is compiled once per part with -DFV3_F32_PART=<0..FUSED32_NPART-1>"
#endif

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

#define FV3_IN_PART(cfg) ((cfg) % FUSED32_NPART == FV3_F32_PART)
#define FV3_CASE(cfg, ...)                                 \
  case cfg:                                                \
    if constexpr (FV3_IN_PART(cfg)) return go<__VA_ARGS__>(a, s); \
    break;

// the configurations of this part; 1 when cfg belongs to another part
#define FV3_PART_FN2(p) launch_fused_f32_part##p
#define FV3_PART_FN(p) FV3_PART_FN2(p)
int FV3_PART_FN(FV3_F32_PART)(const FusedArgs<float> &a, int cfg, const KindStreams &s) {
  switch (cfg) {
    FV3_CASE(0, 4, 2)
    FV3_CASE(1, 8, 1)
    FV3_CASE(2, 4, 1)
    FV3_CASE(3, 2, 2)
    FV3_CASE(4, 6, 1)
    FV3_CASE(5, 1, 2)
    FV3_CASE(6, 1, 1)
    FV3_CASE(7, 2, 1)
    FV3_CASE(8, 4, 1, 3)
    FV3_CASE(9, 4, 2, 3, 4, 1)
    FV3_CASE(10, 4, 1, 3, 4, 1, 4)
    FV3_CASE(11, 4, 2, 3, 4, 1, 3)
    FV3_CASE(12, 8, 2, 3, 4, 1)
    FV3_CASE(13, 6, 2, 3, 4, 1)
    FV3_CASE(14, 8, 1, 3, 4, 1)
    default: break;
  }
  return set_error("fused FP32 launch configuration %d is not instantiated in part %d", cfg, FV3_F32_PART);
}
#undef FV3_CASE

#if FV3_F32_PART == 0
static_assert(FUSED32_NPART == 6, "the dispatcher below names every part");
int launch_fused_f32_part1(const FusedArgs<float> &, int, const KindStreams &);
int launch_fused_f32_part2(const FusedArgs<float> &, int, const KindStreams &);
int launch_fused_f32_part3(const FusedArgs<float> &, int, const KindStreams &);
int launch_fused_f32_part4(const FusedArgs<float> &, int, const KindStreams &);
int launch_fused_f32_part5(const FusedArgs<float> &, int, const KindStreams &);

// W x NT of the lean kind of each case above
static const int kRound[] = {8, 8, 4, 4, 6, 2, 1, 2, 4, 8, 4, 8, 16, 12, 8};
static_assert(sizeof(kRound) / sizeof(kRound[0]) == FUSED32_NCFG, "one entry per launch configuration");

int fused32_round(int cfg) { return cfg >= 0 && cfg < FUSED32_NCFG ? kRound[cfg] : 8; }

int launch_fused_f32(const FusedArgs<float> &a, int cfg, const KindStreams &s) {
  if (cfg < 0 || cfg >= FUSED32_NCFG) return set_error("unknown fused FP32 launch configuration %d", cfg);
  switch (cfg % FUSED32_NPART) {
    case 0: return launch_fused_f32_part0(a, cfg, s);
    case 1: return launch_fused_f32_part1(a, cfg, s);
    case 2: return launch_fused_f32_part2(a, cfg, s);
    case 3: return launch_fused_f32_part3(a, cfg, s);
    case 4: return launch_fused_f32_part4(a, cfg, s);
    default: return launch_fused_f32_part5(a, cfg, s);
  }
}
#endif

}  // namespace fv3
