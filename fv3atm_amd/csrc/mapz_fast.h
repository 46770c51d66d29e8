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

// Remap nq tracers (iv=0, kord_tr) from q_src into q_dst and, when pt_src is
// given, pt (iv=1, kord_tm, t_min) into pt_dst, from source thickness dp1 to
// the ak/bk levels at the source surface pressure; dp2 lands in delp_out.
// Interior columns of one tile; sources and destinations must not overlap.
int remap_fast_f32(Ctx &c, int km, int nq, const double *ak, const double *bk, double ptop,
                   const double *dp1, double *delp_out, const float *q_src, float *q_dst,
                   const float *pt_src, float *pt_dst, int kord_tr, int kord_tm, double t_min,
                   const RemapCols &cols);

}  // namespace fv3
