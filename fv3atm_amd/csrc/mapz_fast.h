// mapz_fast.h -- FP32 column-shared remap data and the fast remap launcher
// used by the fused step (host-side internal interface).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace fv3 {

struct Ctx;

// Column data shared by every field of a column, [level][column] over the
// interior columns (column = j * ni + i):
//   d4, rbt, gg : tridiagonal elimination factors of the interface solve
//                 (level 1 holds the top-row factors)
//   dp1         : source (Lagrangian) layer thickness
//   bot         : [3][column] bottom-row factors
//   lk, cut     : per target layer k, the source layer holding its top
//                 interface, that interface's fraction s within the layer,
//                 the pressure widths of the target's top and bottom pieces,
//                 and 1 / dp2(k)
struct RemapCols {
  float *d4, *rbt, *gg, *dp1, *bot;
  int *lk;
  float4 *cut;
};

size_t remap_cols_bytes(int km, long long ncol);
RemapCols remap_cols_carve(void *base, int km, long long ncol);

// Remap of one tile's interior columns in two launches.  The prep (column
// pressures, Thomas factors, target cut list; dp2 into delp_out) depends on
// the thicknesses only, so it may run before or beside the transport; the
// field kernel then remaps nq tracers (iv=0, kord_tr) from q_src into q_dst
// and, when pt_src is given, pt (iv=1, kord_tm, t_min) into pt_dst.  Sources
// and destinations must not overlap.
int remap_fast_check(int km, long long ncol, long long s2, int kord_tr, int kord_tm, bool with_pt);
// once per process, before the first remap launch: the lane remap kernel's LDS limit
int remap_rows_setup();
int remap_fast_prep_f32(Ctx &c, int km, const double *ak, const double *bk, double ptop,
                        const double *dp1, double *delp_out, const RemapCols &cols, hipStream_t s);
int remap_fast_fields_f32(Ctx &c, int km, int nq, const float *q_src, float *q_dst, const float *pt_src,
                          float *pt_dst, int kord_tr, int kord_tm, double t_min, const RemapCols &cols,
                          hipStream_t s);

}  // namespace fv3
