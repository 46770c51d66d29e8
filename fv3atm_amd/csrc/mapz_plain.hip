// mapz_plain.hip -- the plain form of the tracer/pt remap of
// Lagrangian_to_Eulerian (upstream fv_mapz.F90: mapn_tracer, map1_q2,
// map1_ppm, map_scalar, scalar_profile / cs_profile, cs_limiters, fillz).
//
// One thread owns one column; its per-level work arrays live in a global
// scratch laid out [level][column] so that each level access of a wavefront
// is one coalesced 64-column read.  Pressure interfaces, layer thicknesses and
// the source-layer search are FP64; the profile arithmetic is in T.
#include <algorithm>

#include "fv3_internal.h"
#include "mapz_column.h"

namespace fv3 {

// Column kernel.  mode 0: Lagrangian_to_Eulerian tracer subset (pe1 from
// delp, pe2 from ak/bk, every tracer then pt, delp <- dp2).  mode 1: one
// field with given pe1/pe2 (map1_ppm / map_scalar / map1_q2 rows).
template <typename T>
__global__ void __launch_bounds__(64) k_remap_cols(RemapArgs<T> a) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  remap_column<T>(a, col);
}

template <typename T>
int remap_launch(Ctx &c, RemapArgs<T> &a) {
  const int km = a.km, kn = a.kn;
  const long long ncol = a.ncol;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  const int kmax = std::max(km, kn);
  const size_t o_pe1 = take(sizeof(double) * (km + 1) * ncol);
  const size_t o_pe2 = take(sizeof(double) * (kn + 1) * ncol);
  const size_t o_dp1 = take(sizeof(double) * km * ncol);
  const size_t o_dp2 = take(sizeof(double) * kn * ncol);
  const size_t o_a4 = take(sizeof(T) * 4 * (size_t)km * ncol);
  const size_t o_qi = take(sizeof(T) * (km + 1) * ncol);
  const size_t o_gam = take(sizeof(T) * (km + 1) * ncol);
  const size_t o_flg = take(sizeof(uint8_t) * 3 * (size_t)km * ncol);
  const size_t o_dm = take(sizeof(T) * (kmax + 1) * ncol);
  char *base = (char *)ctx_scratch(c, off);
  if (!base) return -1;
  a.pe1 = (double *)(base + o_pe1);
  a.pe2 = (double *)(base + o_pe2);
  a.dp1 = (double *)(base + o_dp1);
  a.dp2 = (double *)(base + o_dp2);
  a.a4 = (T *)(base + o_a4);
  a.qi = (T *)(base + o_qi);
  a.gam = (T *)(base + o_gam);
  a.flags = (uint8_t *)(base + o_flg);
  a.dm = (T *)(base + o_dm);
  a.err = nullptr;
  dim3 b(64), g((unsigned)((ncol + 63) / 64));
  hipLaunchKernelGGL(k_remap_cols<T>, g, b, 0, c.stream, a);
  FV3_HIP(hipGetLastError());
  FV3_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

template <typename T>
int remap_tracers_plain(Ctx &c, int km, int nq, const double *ak, const double *bk, double ptop,
                        double *delp, T *q, T *pt, const int *kord_tr, int kord_tm, double t_min,
                        int fill) {
  if (km < 4) return set_error("remap needs km >= 4 (got %d)", km);
  for (int iq = 0; iq < nq; iq++)
    if (!(kord_tr[iq] > 7 && (kord_tr[iq] <= 10 || kord_tr[iq] > 16)))
      return set_error("kord_tr(%d)=%d: implemented orders are 8, 9, 10 and > 16", iq + 1, kord_tr[iq]);
  if (pt && !(kord_tm > 7 && (kord_tm <= 10 || kord_tm > 16)))
    return set_error("kord_tm=%d: implemented orders are 8, 9, 10 and > 16", kord_tm);
  if (nq > REMAP_MAX_NQ) return set_error("nq=%d exceeds %d", nq, REMAP_MAX_NQ);
  const Dims &d = c.d;
  RemapArgs<T> a{};
  a.mode = 0;
  a.km = km;
  a.kn = km;
  a.nq = nq;
  a.ntiles = c.ntiles;
  a.ni = d.ie - d.is + 1;
  a.nj = d.je - d.js + 1;
  a.ncol = (long long)a.ni * a.nj * c.ntiles;
  a.is = d.is;
  a.js = d.js;
  a.isd = d.isd;
  a.jsd = d.jsd;
  a.nid = d.nid;
  a.s2 = d.s2;
  a.ak = ak;
  a.bk = bk;
  a.ptop = ptop;
  a.delp = delp;
  a.q = q;
  a.pt = pt;
  for (int iq = 0; iq < nq; iq++) a.kord[iq] = kord_tr[iq];
  a.kord_tm = kord_tm;
  a.t_min = (T)t_min;
  a.fill = fill;
  a.mapn = nq > 5;
  return remap_launch<T>(c, a);
}

template int remap_tracers_plain<double>(Ctx &, int, int, const double *, const double *, double,
                                         double *, double *, double *, const int *, int, double, int);
template int remap_tracers_plain<float>(Ctx &, int, int, const double *, const double *, double,
                                        double *, float *, float *, const int *, int, double, int);
template int remap_launch<double>(Ctx &, RemapArgs<double> &);
template int remap_launch<float>(Ctx &, RemapArgs<float> &);

__global__ void __launch_bounds__(64) k_cs_rows(const double *a1, const double *dp, const double *qs,
                                                double *out, int im, int km, int iv, int kord,
                                                double *qi, double *gam, uint8_t *flags) {
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= im) return;
  const long long n = (long long)km * im;
  CV<double> A1{out + col, im}, A2{out + n + col, im}, A3{out + 2 * n + col, im}, A4{out + 3 * n + col, im};
  for (int k = 1; k <= km; k++) A1(k) = a1[(long long)(k - 1) * im + col];
  CV<double> DP{const_cast<double *>(dp) + col, im}, Q{qi + col, im}, G{gam + col, im};
  CV<uint8_t> E1{flags + col, im}, E5{flags + n + col, im}, E6{flags + 2 * n + col, im};
  cs_profile_col<double>(A1, A2, A3, A4, DP, qs ? qs[col] : 0.0, km, iv, kord, false, 0.0, Q, G, E1, E5, E6);
}

int cs_profile_rows(Ctx &c, const double *a1, const double *dp, const double *qs, double *out,
                    int im, int km, int iv, int kord) {
  const size_t nq = (size_t)(km + 1) * im;
  char *base = (char *)ctx_scratch(c, sizeof(double) * 2 * nq + 3 * (size_t)km * im + 1024);
  if (!base) return -1;
  double *qi = (double *)base, *gam = qi + nq;
  uint8_t *flags = (uint8_t *)(gam + nq);
  hipLaunchKernelGGL(k_cs_rows, dim3((im + 63) / 64), dim3(64), 0, c.stream, a1, dp, qs, out, im, km,
                     iv, kord, qi, gam, flags);
  FV3_HIP(hipGetLastError());
  FV3_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

}  // namespace fv3
