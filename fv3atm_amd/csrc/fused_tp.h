// fused_tp.h -- one-kernel-per-sub-cycle form of fv_tp_2d + the tracer_2d
// flux-form update (upstream tp_core.F90 / fv_tracer2d.F90) for MI355X.
//
// Mapping.  A wavefront owns a strip of 64 columns (58 interior columns plus
// the 3-cell halo on each side) and streams down the rows of a block of H
// interior rows.  At each input row r it
//   - slides the y-direction PPM window (registers only) to get fy2 at face
//     r-2 and the transverse-updated q_i of row r-3,
//   - sweeps row r-3 of q_i in x across lanes (shuffles) for fx,
//   - sweeps row r of q in x for fx2 and forms q_j of row r,
//   - slides the y-direction window over q_j for fy at face r-2,
//   - averages the fluxes and writes the flux-form update of row r-3.
// Nothing but q enters or leaves HBM per tracer; the winds, mass fluxes and
// thicknesses of the level are shared by the waves of a workgroup (each wave
// one tracer, the waves kept in step by a barrier per row, so the level data
// is fetched once per workgroup).  The 1/nsplit scaling of cx, cy, mfx, mfy
// and the dp1 <- dp2 chain of sub-cycling are applied in flight.
//
// Every expression keeps the upstream evaluation order: the T = double
// instantiation, built with contraction off, reproduces the CPU FP64
// evaluation; T = float is the bulk FP32 form (pressure thickness in FP64).
#pragma once
#include <hip/hip_runtime.h>

#include "fv3_device.h"

namespace fv3 {

constexpr int FT_LANES = 64;
constexpr int FT_BX = FT_LANES - 2 * NG;  // interior columns per strip

template <typename T>
struct FusedArgs {
  // geometry (one tile)
  int is, ie, js, je, isd, ied, jsd, jed, npx, npy;
  int nid;
  long long s2, sx, sy, smx, smy;
  unsigned corners;
  bool edge;
  int npz, nq, hord, it, cap, H, nstrip, nvblk;
  const int *nsplit;  // per level
  // fields (device)
  const T *q_in;  // step input (also the frozen halo source)
  T *q_out, *q_tmp;
  const T *cx, *cy, *mfx, *mfy;
  const double *dp1;
  // metrics (one tile)
  const T *area, *rarea, *dxa, *dya, *dx, *dy, *sin_sg;
  const T *gxp, *gxm, *gyp, *gym;  // FP32 form: combined upwind area-flux factors
  long long ms2;                   // metric slab (isd:ied, jsd:jed)
};

template <typename T>
__device__ __forceinline__ T shup(T v, int d) {
  return __shfl_up(v, d, FT_LANES);
}
template <typename T>
__device__ __forceinline__ T shdn(T v, int d) {
  return __shfl_down(v, d, FT_LANES);
}

// Lin's monotone slope from three cells
template <typename T>
__device__ __forceinline__ T dm3(T qm, T q0, T qp) {
  T xt = T(0.25) * (qp - qm);
  return fsign(mn3(fabs(xt), mx3(qm, q0, qp) - q0, q0 - mn3(qm, q0, qp)), xt);
}
// face value between cells m-1 and m
template <typename T>
__device__ __forceinline__ T al2(T qm, T q0, T dmm, T dm0) {
  return T(0.5) * (qm + q0) + K<T>::r3 * (dmm - dm0);
}
// cube-edge face value between b and c from cells a b c e with spacings
template <typename T>
__device__ __forceinline__ T edge_face(T a, T b, T c, T e, T da, T db, T dc, T de) {
  T xt = T(0.5) * (((T(2) * db + da) * b - db * a) / (da + db) + ((T(2) * dc + de) * c - dc * e) / (dc + de));
  xt = mx2(xt, mn4(a, b, c, e));
  xt = mn2(xt, mx4(a, b, c, e));
  return xt;
}

// Cube-edge role of cell m along a line: 0 none, 1..3 cells 0,1,2, 4..6 cells
// npx-2, npx-1, npx.
__device__ __forceinline__ int edge_role(int m, int is, int ie, int npx, bool edge) {
  if (!edge) return 0;
  if (is == 1 && m >= 0 && m <= 2) return 1 + m;
  if (ie + 1 == npx && m >= npx - 2 && m <= npx) return 4 + (m - (npx - 2));
  return 0;
}

// (bl, br) of one cell from its neighbourhood (xppm/yppm cell logic).
// d0..d4: spacings of cells m-2..m+2 (read only for edge roles).
template <typename T>
__device__ __forceinline__ void cell_coef(T qm2, T qm1, T q0, T qp1, T qp2, T dmm1, T dm0, T dmp1,
                                          T al0, T alp1, int role, int iord, T d0, T d1, T d2,
                                          T d3, T d4, T &bl, T &br) {
  if (role != 0) {
    switch (role) {
      case 1:
        bl = K<T>::s14 * dmm1 + K<T>::s11 * (qm1 - q0);
        br = edge_face(qm1, q0, qp1, qp2, d1, d2, d3, d4) - q0;
        break;
      case 2:
        bl = edge_face(qm2, qm1, q0, qp1, d0, d1, d2, d3) - q0;
        br = K<T>::s15 * q0 + K<T>::s11 * qp1 - K<T>::s14 * dmp1 - q0;
        break;
      case 3:
        bl = K<T>::s15 * qm1 + K<T>::s11 * q0 - K<T>::s14 * dm0 - q0;
        br = alp1 - q0;
        break;
      case 4:
        bl = al0 - q0;
        br = K<T>::s15 * qp1 + K<T>::s11 * q0 + K<T>::s14 * dm0 - q0;
        break;
      case 5:
        bl = K<T>::s15 * q0 + K<T>::s11 * qm1 + K<T>::s14 * dmm1 - q0;
        br = edge_face(qm1, q0, qp1, qp2, d1, d2, d3, d4) - q0;
        break;
      default:
        bl = edge_face(qm2, qm1, q0, qp1, d0, d1, d2, d3) - q0;
        br = K<T>::s11 * (qp1 - q0) - K<T>::s14 * dmp1;
        break;
    }
    pert_ppm1(q0, bl, br, 1);
    return;
  }
  if (iord == 8 || iord == 11) {
    T xt = iord == 8 ? T(2) * dm0 : K<T>::ppm_fac * dm0;
    bl = -fsign(mn2(fabs(xt), fabs(al0 - q0)), xt);
    br = fsign(mn2(fabs(xt), fabs(alp1 - q0)), xt);
  } else {
    bl = al0 - q0;
    br = alp1 - q0;
    if (fabs(dmm1) + fabs(dm0) + fabs(dmp1) < K<T>::near_zero) {
      bl = T(0);
      br = T(0);
    } else if (fabs(T(3) * (bl + br)) > fabs(bl - br)) {
      T dqm2 = T(2) * (qm1 - qm2), dqm1 = T(2) * (q0 - qm1), dq0 = T(2) * (qp1 - q0), dqp1 = T(2) * (qp2 - qp1);
      T pmp_2 = dqm1;
      T lac_2 = pmp_2 - T(0.75) * dqm2;
      br = mn2(mx3(T(0), pmp_2, lac_2), mx2(br, mn3(T(0), pmp_2, lac_2)));
      T pmp_1 = -dq0;
      T lac_1 = pmp_1 + T(0.75) * dqp1;
      bl = mn2(mx3(T(0), pmp_1, lac_1), mx2(bl, mn3(T(0), pmp_1, lac_1)));
    }
    if (iord == 9 || iord == 13) pert_ppm1(q0, bl, br, 0);
  }
}

// upwind flux at a face from the upwind cell's (q, bl, br)
template <typename T>
__device__ __forceinline__ T face_flux(T c, T qu, T blu, T bru) {
  if (c > T(0)) return qu + (T(1) - c) * (bru - c * (blu + bru));
  return qu + (T(1) + c) * (blu + c * (blu + bru));
}

// x-direction PPM across the lanes of a row: flux at the face left of each
// lane's cell.  dxl: dxa of the lane's cell (used in edge strips only).
template <typename T>
__device__ __forceinline__ T xsweep(T q, T c, int i, int is, int ie, int npx, bool edge,
                                    bool edge_strip, int iord, T dxl) {
  const T qm1 = shup(q, 1), qp1 = shdn(q, 1);
  const T dm = dm3(qm1, q, qp1);
  const T dmm1 = shup(dm, 1), dmp1 = shdn(dm, 1);
  const T al = al2(qm1, q, dmm1, dm);
  const T alp1 = shdn(al, 1);
  T qm2 = q, qp2 = q;
  if (iord != 8 || edge_strip) {
    qm2 = shup(q, 2);
    qp2 = shdn(q, 2);
  }
  T d0 = dxl, d1 = dxl, d3 = dxl, d4 = dxl;
  int role = 0;
  if (edge_strip) {
    d0 = shup(dxl, 2);
    d1 = shup(dxl, 1);
    d3 = shdn(dxl, 1);
    d4 = shdn(dxl, 2);
    role = edge_role(i, is, ie, npx, edge);
  }
  T bl, br;
  cell_coef(qm2, qm1, q, qp1, qp2, dmm1, dm, dmp1, al, alp1, role, iord, d0, d1, dxl, d3, d4, bl, br);
  const T blm = shup(bl, 1), brm = shup(br, 1);
  return c > T(0) ? face_flux(c, qm1, blm, brm) : face_flux(c, q, bl, br);
}

// Sliding y-direction PPM over a column: fed one cell per row, yields the
// flux at face r-2 once rows r-5..r are in.  Histories are indexed by the
// row modulo 6 so that an unrolled loop keeps them in fixed registers.
template <typename T>
struct YPipe {
  T w[6];   // q of rows r-5..r, slot row%6
  T dm[6];  // dm of cells, slot cell%6
  T al[6];  // face values, slot face%6
  T bl[6], br[6];
};

// one step at input row r; U is the slot of row r (the step index modulo 6).
template <typename T, int U>
__device__ __forceinline__ T ypipe_step(YPipe<T> &p, T qnew, int r, T c, int js, int je, int npy,
                                        bool edge, int iord, const T *dcol, long long dstride,
                                        int jsd, int jed) {
  constexpr int s0 = U % 6, s1 = (U + 5) % 6, s2 = (U + 4) % 6, s3 = (U + 3) % 6,
                s4 = (U + 2) % 6;  // rows r, r-1, r-2, r-3, r-4
  p.w[s0] = qnew;
  // dm of cell r-1, face value al at face r-1 (between cells r-2 and r-1)
  p.dm[s1] = dm3(p.w[s2], p.w[s1], p.w[s0]);
  p.al[s1] = al2(p.w[s2], p.w[s1], p.dm[s2], p.dm[s1]);
  // coefficients of cell m = r-2
  const int m = r - 2;
  const int role = edge_role(m, js, je, npy, edge);
  T d0 = T(0), d1 = T(0), d2 = T(0), d3 = T(0), d4 = T(0);
  if (role != 0) {  // uniform across the wave (all lanes share the row)
    auto dd = [&](int jj) { jj = jj < jsd ? jsd : (jj > jed ? jed : jj); return dcol[(long long)(jj - jsd) * dstride]; };
    d0 = dd(m - 2);
    d1 = dd(m - 1);
    d2 = dd(m);
    d3 = dd(m + 1);
    d4 = dd(m + 2);
  }
  cell_coef(p.w[s4], p.w[s3], p.w[s2], p.w[s1], p.w[s0], p.dm[s3], p.dm[s2], p.dm[s1], p.al[s2],
            p.al[s1], role, iord, d0, d1, d2, d3, d4, p.bl[s2], p.br[s2]);
  // face r-2: upwind cell r-3 (c > 0) or r-2
  return c > T(0) ? face_flux(c, p.w[s3], p.bl[s3], p.br[s3]) : face_flux(c, p.w[s2], p.bl[s2], p.br[s2]);
}

}  // namespace fv3
