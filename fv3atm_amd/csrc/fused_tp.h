// fused_tp.h -- one-kernel-per-sub-cycle form of fv_tp_2d + the tracer_2d
// flux-form update (upstream tp_core.F90 / fv_tracer2d.F90) for MI355X.
//
// Mapping.  A wavefront owns a strip of 64 columns (58 interior columns plus
// the 3-cell halo on each side) and streams down the rows of a block of H
// interior rows, carrying NT tracers of one level at once.  At each input
// row r it
//   - slides the y-direction PPM window (registers only) to get fy2 at face
//     r-2 and the transverse-updated q_i of row r-3,
//   - sweeps row r-3 of q_i in x across lanes (DPP wave shifts) for fx,
//   - sweeps row r of q in x for fx2 and forms q_j of row r,
//   - slides the y-direction window over q_j for fy at face r-2,
//   - averages the fluxes and writes the flux-form update of row r-3.
// Nothing but q enters or leaves HBM per tracer.  The winds, mass fluxes and
// thicknesses of the level are read once per row for the NT tracers of a
// wave, and the waves of a workgroup (all on the same level and strip) are
// kept in step by a barrier per row so those reads hit L1.  The 1/nsplit
// scaling of cx, cy, mfx, mfy and the dp1 <- dp2 chain of sub-cycling are
// applied in flight.
//
// The T = double instantiation, built with contraction off, keeps the
// upstream evaluation order and reproduces the CPU FP64 evaluation; T = float
// is the bulk FP32 form (reciprocal multiplies, contraction on).
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>

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
  // Interior strips [s_lo, s_hi] touch no cube-edge column or halo lane.
  // Their rows lt..lb touch no cube-edge row or halo row either: the lean
  // kernel variant takes them in nmid blocks of H rows; the edge-row variant
  // takes rows js..lt-1 and lb+1..je.  Every other strip is cut into nvblk
  // blocks of H rows.
  int s_lo, s_hi, lt, lb, nmid;
  // six tiles (blockIdx.z = tile): per-tile strides of q, winds, dp1, the
  // metrics and the FP32 factors; halo cells are read from the neighbour
  // tile's interior in the pass input buffer
  int ntiles;
  long long tq, tcx, tcy, tmx, tmy, tdp, tgs2, tgdx, tgdy, tgf;
  const int *nsplit;  // per level
  // fields (device)
  const T *q_in;  // step input (also the frozen halo source)
  T *q_out, *q_tmp;
  const T *cx, *cy, *mfx, *mfy;
  const double *dp1;
  // metrics (one tile)
  const T *area, *rarea, *dxa, *dya, *dx, *dy, *sin_sg;
  const T *gxp, *gxm, *gyp, *gym;  // FP32 form: combined upwind area-flux factors
  const T *gex;  // FP32 form: per-cell weights of the cube-edge face value (tile stride tgf)
  long long ms2;                   // metric slab (isd:ied, jsd:jed)
};

// Block counts of the kernel kinds over (strip, row block): n[0] lean
// interior blocks, n[1] edge-row blocks of the interior strips, n[2] edge
// strips when there is no interior strip, n[3] / n[4] the strips left / right
// of the interior strips.
template <typename T>
inline void fused_grid_split(const FusedArgs<T> &a, int (&n)[7]) {
  const int ns = a.s_hi >= a.s_lo ? a.s_hi - a.s_lo + 1 : 0;
  const int nl = ns > 0 ? a.s_lo : 0, nr = ns > 0 ? a.nstrip - 1 - a.s_hi : 0;
  n[0] = ns * a.nmid;
  n[1] = ns * 2;  // the top and bottom edge-row blocks of each interior strip
  n[2] = ns > 0 ? 0 : a.nstrip * a.nvblk;
  n[3] = nl * 2;  // the strips left of the interior strips: top and bottom blocks
  n[4] = nr * 2;  // right of them: top and bottom blocks
  n[5] = nl * a.nmid;  // left strips, lean rows
  n[6] = nr * a.nmid;  // right strips, lean rows
}

// ---- cross-lane shifts: value of lane l-1 / l+1 (DPP wave_shr:1 / wave_shl:1)
// (bound_ctrl: the lane past the wave edge reads 0, so no destination
// initialisation is needed; that lane is a halo lane whose result is unused)
__device__ __forceinline__ int dpp_from_left(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true); }
__device__ __forceinline__ int dpp_from_right(int v) { return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true); }

template <typename T>
__device__ __forceinline__ T from_left(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(dpp_from_left(__float_as_int(v)));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_left((int)(b & 0xffffffffll)), hi = dpp_from_left((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}
template <typename T>
__device__ __forceinline__ T from_right(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(dpp_from_right(__float_as_int(v)));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_right((int)(b & 0xffffffffll)), hi = dpp_from_right((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// min/max as single instructions (max3/min3 fused); equal in value to the
// upstream MIN/MAX on every finite input
template <typename T> __device__ __forceinline__ T fmx(T a, T b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T fmn(T a, T b) { return fmin(a, b); }
template <typename T> __device__ __forceinline__ T fmx3(T a, T b, T c) { return fmax(fmax(a, b), c); }
template <typename T> __device__ __forceinline__ T fmn3(T a, T b, T c) { return fmin(fmin(a, b), c); }

// x / y: exact for FP64, reciprocal multiply for FP32
template <typename T>
__device__ __forceinline__ T divide(T x, T y) {
  if constexpr (sizeof(T) == 4)
    return x * __builtin_amdgcn_rcpf(y);
  else
    return x / y;
}

// Lin's monotone slope from three cells
template <typename T>
__device__ __forceinline__ T dm3(T qm, T q0, T qp) {
  T xt = T(0.25) * (qp - qm);
  return fsign(fmn3(fabs(xt), fmx3(qm, q0, qp) - q0, q0 - fmn3(qm, q0, qp)), xt);
}
// face value between cells m-1 and m
template <typename T>
__device__ __forceinline__ T al2(T qm, T q0, T dmm, T dm0) {
  return T(0.5) * (qm + q0) + K<T>::r3 * (dmm - dm0);
}
// cube-edge face value between b and c from cells a b c e with spacings
// da..de (x / y: exact division for FP64, reciprocal multiply for FP32)
template <typename T>
__device__ __forceinline__ T edge_face(T a, T b, T c, T e, T da, T db, T dc, T de) {
  T xt = T(0.5) * (divide((T(2) * db + da) * b - db * a, da + db) + divide((T(2) * dc + de) * c - dc * e, dc + de));
  xt = fmx(xt, fmn(fmn3(a, b, c), e));
  xt = fmn(xt, fmx(fmx3(a, b, c), e));
  return xt;
}

// Cube-sphere edge halo: the source (tile ts, cell si, sj; 1-based tiles) of
// edge-halo cell (i, j) of tile t on a cube of n x n tiles, for the FMS
// mosaic's face orientation (odd and even faces alternate orientation).
// Exactly one of i, j lies outside 1..n.
__device__ __host__ __forceinline__ void cube_halo_src(int t, int i, int j, int n, int &ts, int &si, int &sj) {
  auto nxt = [](int tt, int s) { return ((tt - 1 + s) % 6 + 6) % 6 + 1; };
  const bool odd = (t & 1) != 0;
  if (i > n) {  // east, depth e, along a
    const int e = i - n, a = j;
    if (odd) { ts = nxt(t, 1); si = e; sj = a; }
    else { ts = nxt(t, 2); si = n + 1 - a; sj = e; }
  } else if (j > n) {  // north
    const int e = j - n, a = i;
    if (odd) { ts = nxt(t, 2); si = e; sj = n + 1 - a; }
    else { ts = nxt(t, 1); si = a; sj = e; }
  } else if (i < 1) {  // west
    const int e = 1 - i, a = j;
    if (odd) { ts = nxt(t, -2); si = n + 1 - a; sj = n + 1 - e; }
    else { ts = nxt(t, -1); si = n + 1 - e; sj = a; }
  } else {  // south
    const int e = 1 - j, a = i;
    if (odd) { ts = nxt(t, -1); si = a; sj = n + 1 - e; }
    else { ts = nxt(t, -2); si = n + 1 - e; sj = n + 1 - a; }
  }
}

// Cube-edge role of cell m along a line: 0 none, 1..3 cells 0,1,2, 4..6 cells
// npx-2, npx-1, npx.
__device__ __forceinline__ int edge_role(int m, int is, int ie, int npx, bool edge) {
  if (!edge) return 0;
  if (is == 1 && m >= 0 && m <= 2) return 1 + m;
  if (ie + 1 == npx && m >= npx - 2 && m <= npx) return 4 + (m - (npx - 2));
  return 0;
}

// (bl, br) of a cell with a cube-edge role (xppm/yppm edge code): role
// 1..3 are cells 0, 1, 2 of the line, 4..6 cells npx-2, npx-1, npx.  Every
// candidate is computed and the role selects (no divergent branches across
// the lanes of an x-sweep); role 0 keeps the interior (bl, br).  efl / efr:
// the edge face values on the cell's left / right.
template <typename T>
__device__ __forceinline__ void edge_coef_sel(T qm1, T q0, T qp1, T dmm1, T dm0, T dmp1, T al0, T alp1,
                                              int role, T efl, T efr, T &bl, T &br) {
  const T b1 = K<T>::s14 * dmm1 + K<T>::s11 * (qm1 - q0);
  const T b3 = K<T>::s15 * qm1 + K<T>::s11 * q0 - K<T>::s14 * dm0 - q0;
  const T b4 = al0 - q0;
  const T b5 = K<T>::s15 * q0 + K<T>::s11 * qm1 + K<T>::s14 * dmm1 - q0;
  const T bL = efl - q0;  // roles 2, 6
  const T r2 = K<T>::s15 * q0 + K<T>::s11 * qp1 - K<T>::s14 * dmp1 - q0;
  const T r3 = alp1 - q0;
  const T r4 = K<T>::s15 * qp1 + K<T>::s11 * q0 + K<T>::s14 * dm0 - q0;
  const T r6 = K<T>::s11 * (qp1 - q0) - K<T>::s14 * dmp1;
  const T rR = efr - q0;  // roles 1, 5
  T l = role == 1 ? b1 : (role == 3 ? b3 : (role == 4 ? b4 : (role == 5 ? b5 : bL)));
  T r = role == 2 ? r2 : (role == 3 ? r3 : (role == 4 ? r4 : (role == 6 ? r6 : rR)));
  pert_ppm1(q0, l, r, 1);
  if (role != 0) {
    bl = l;
    br = r;
  }
}

// The same for a strip that can only touch the left cube edge (roles 1..3,
// SIDE 1) or only the right one (roles 4..6, SIDE 2): half the candidates.
template <int SIDE, typename T>
__device__ __forceinline__ void edge_coef_side(T qm1, T q0, T qp1, T dmm1, T dm0, T dmp1, T al0, T alp1,
                                               int role, T efl, T efr, T &bl, T &br) {
  T l, r;
  if constexpr (SIDE == 1) {
    const T b1 = K<T>::s14 * dmm1 + K<T>::s11 * (qm1 - q0);
    const T b3 = K<T>::s15 * qm1 + K<T>::s11 * q0 - K<T>::s14 * dm0 - q0;
    const T bL = efl - q0;  // role 2
    const T r2 = K<T>::s15 * q0 + K<T>::s11 * qp1 - K<T>::s14 * dmp1 - q0;
    const T r3 = alp1 - q0;
    const T rR = efr - q0;  // role 1
    l = role == 1 ? b1 : (role == 3 ? b3 : bL);
    r = role == 2 ? r2 : (role == 3 ? r3 : rR);
  } else {
    const T b4 = al0 - q0;
    const T b5 = K<T>::s15 * q0 + K<T>::s11 * qm1 + K<T>::s14 * dmm1 - q0;
    const T bL = efl - q0;  // role 6
    const T r4 = K<T>::s15 * qp1 + K<T>::s11 * q0 + K<T>::s14 * dm0 - q0;
    const T r6 = K<T>::s11 * (qp1 - q0) - K<T>::s14 * dmp1;
    const T rR = efr - q0;  // role 5
    l = role == 4 ? b4 : (role == 5 ? b5 : bL);
    r = role == 4 ? r4 : (role == 6 ? r6 : rR);
  }
  pert_ppm1(q0, l, r, 1);
  if (role != 0) {
    bl = l;
    br = r;
  }
}

// pert_ppm's standard monotone constraint (iv = 1) with selects instead of
// branches: each outcome is the same expression as pert_ppm1's, so the values
// are identical; an edge strip evaluates it on every lane of every x-sweep
__device__ __forceinline__ void pert_ppm1_sel(float &al, float &ar) {
  const float da1 = al - ar, da2 = da1 * da1, a6da = 3.f * (al + ar) * da1;
  const bool ext = al * ar < 0.f;
  const float nl = a6da > da2 ? -2.f * ar : al;
  const float nr = a6da < -da2 ? -2.f * al : ar;
  al = ext ? nl : 0.f;
  ar = ext ? nr : 0.f;
}

// FP32 edge strips: a lane's cube-edge (bl, br) candidates as fixed linear
// combinations of its neighbourhood,
//   l = cl . (qm1, q0, dmm1, dm0, al0, efl),  r = cr . (q0, qp1, dmp1, dm0, alp1, efr),
// with the coefficients of its edge role (the same sums as edge_coef_sel);
// on: the lane has a role (its interior (bl, br) are replaced).
struct EdgeCoef {
  float l[6], r[6];
  bool on;
};

__device__ __forceinline__ EdgeCoef edge_coef_of(int role) {
  constexpr float s11 = K<float>::s11, s14 = K<float>::s14, s15 = K<float>::s15;
  EdgeCoef e;
#pragma unroll
  for (int m = 0; m < 6; m++) e.l[m] = e.r[m] = 0.f;
  e.on = role != 0;
  switch (role) {
    case 1:  // cell 0: s14 dm(-1) + s11 (q(-1) - q0); efr - q0
      e.l[0] = s11; e.l[1] = -s11; e.l[2] = s14;
      e.r[0] = -1.f; e.r[5] = 1.f;
      break;
    case 2:  // cell 1: efl - q0; s15 q0 + s11 q(+1) - s14 dm(+1) - q0
      e.l[1] = -1.f; e.l[5] = 1.f;
      e.r[0] = s15 - 1.f; e.r[1] = s11; e.r[2] = -s14;
      break;
    case 3:  // cell 2: s15 q(-1) + s11 q0 - s14 dm0 - q0; al(+1) - q0
      e.l[0] = s15; e.l[1] = s11 - 1.f; e.l[3] = -s14;
      e.r[0] = -1.f; e.r[4] = 1.f;
      break;
    case 4:  // cell npx-2: al0 - q0; s15 q(+1) + s11 q0 + s14 dm0 - q0
      e.l[1] = -1.f; e.l[4] = 1.f;
      e.r[0] = s11 - 1.f; e.r[1] = s15; e.r[3] = s14;
      break;
    case 5:  // cell npx-1: s15 q0 + s11 q(-1) + s14 dm(-1) - q0; efr - q0
      e.l[0] = s11; e.l[1] = s15 - 1.f; e.l[2] = s14;
      e.r[0] = -1.f; e.r[5] = 1.f;
      break;
    case 6:  // cell npx: efl - q0; s11 (q(+1) - q0) - s14 dm(+1)
      e.l[1] = -1.f; e.l[5] = 1.f;
      e.r[0] = -s11; e.r[1] = s11; e.r[2] = -s14;
      break;
    default:
      break;
  }
  return e;
}

// (bl, br) of an interior cell for the monotone orders
template <typename T>
__device__ __forceinline__ void cell_coef(T qm2, T qm1, T q0, T qp1, T qp2, T dmm1, T dm0, T dmp1,
                                          T al0, T alp1, int iord, T &bl, T &br) {
  if (iord == 8 || iord == 11) {
    T xt = iord == 8 ? T(2) * dm0 : K<T>::ppm_fac * dm0;
    if constexpr (sizeof(T) == 4) {
      // -sign(xt) min(|xt|, |b|) is -xt clamped to [-|b|, |b|] (the same
      // value, signed zeros aside): one median-of-three per side
      const T b = al0 - q0, c = alp1 - q0;
      bl = __builtin_amdgcn_fmed3f(-xt, -fabs(b), fabs(b));
      br = __builtin_amdgcn_fmed3f(xt, -fabs(c), fabs(c));
    } else {
      bl = -fsign(fmn(fabs(xt), fabs(al0 - q0)), xt);
      br = fsign(fmn(fabs(xt), fabs(alp1 - q0)), xt);
    }
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
      br = fmn(fmx3(T(0), pmp_2, lac_2), fmx(br, fmn3(T(0), pmp_2, lac_2)));
      T pmp_1 = -dq0;
      T lac_1 = pmp_1 + T(0.75) * dqp1;
      bl = fmn(fmx3(T(0), pmp_1, lac_1), fmx(bl, fmn3(T(0), pmp_1, lac_1)));
    }
    if (iord == 9 || iord == 13) pert_ppm1(q0, bl, br, 0);
  }
}

// Upwind flux at a face with Courant number c, from the cells on its left
// (qL, brL and sL = blL + brL) and right (qR, blR and sR = blR + brR).
// Upstream writes the two branches
//   c > 0:  qL + (1 - c) (brL - c (blL + brL))
//   else:   qR + (1 + c) (blR + c (blR + brR));
// with a = |c| both are  qu + (1 - a) (x - a su)  where u is the upwind cell,
// x its downwind-side perturbation and su its sum: 1 - a equals 1 + c and
// a S equals -(c S) exactly, so this branch-free form rounds the same, and a
// cell's sum bl + br is formed once for the two faces that may use it.
template <typename T>
__device__ __forceinline__ T face_flux_s(T c, T qL, T brL, T sL, T qR, T blR, T sR) {
  const bool pos = c > T(0);
  const T qu = pos ? qL : qR, su = pos ? sL : sR, x = pos ? brL : blR;
  const T ac = fabs(c);
  return qu + (T(1) - ac) * (x - ac * su);
}

// x-direction PPM across the lanes of a row: flux at the face left of each
// lane's cell.  dxl: dxa of the lane's cell (used in edge strips only).
// SIDE: 0 a strip that may touch either cube edge, 1 only the left one, 2
// only the right one (the edge code then evaluates half the candidates).
template <int SIDE = 0, typename T>
__device__ __forceinline__ T xsweep(T q, T c, int i, int is, int ie, int npx, bool edge,
                                    bool edge_strip, int iord, T dxl, const EdgeCoef &ec) {
  const T qm1 = from_left(q), qp1 = from_right(q);
  const T dm = dm3(qm1, q, qp1);
  const T dmm1 = from_left(dm), dmp1 = from_right(dm);
  const T al = al2(qm1, q, dmm1, dm);
  const T alp1 = from_right(al);
  T qm2 = q, qp2 = q;
  if (iord != 8 || edge_strip) {
    qm2 = from_left(qm1);
    qp2 = from_right(qp1);
  }
  T bl, br;
  cell_coef(qm2, qm1, q, qp1, qp2, dmm1, dm, dmp1, al, alp1, iord, bl, br);
  if (edge_strip) {
    if constexpr (sizeof(T) == 4) {
      // dxl: this cell's weight in the edge face value at the right face of
      // the cell left of the cube edge (non-zero for the two cells either
      // side of that face and their outer neighbours): the weighted sum of
      // the four cells starting one to the left, then clipped to their range
      const T p = dxl * q;
      const T pp1 = from_right(p);
      T efr = (from_left(p) + p) + (pp1 + from_right(pp1));
      efr = fmx(efr, fmn(fmn3(qm1, q, qp1), qp2));
      efr = fmn(efr, fmx(fmx3(qm1, q, qp1), qp2));
      const T efl = from_left(efr);
      // roles 1..3 never need al0 (l) or dm0 (r), roles 4..6 never dm0 (l) or alp1 (r)
      T l = ec.l[0] * qm1 + ec.l[1] * q + ec.l[2] * dmm1 + ec.l[5] * efl;
      T r = ec.r[0] * q + ec.r[1] * qp1 + ec.r[2] * dmp1 + ec.r[5] * efr;
      if constexpr (SIDE != 2) {
        l += ec.l[3] * dm;
        r += ec.r[4] * alp1;
      }
      if constexpr (SIDE != 1) {
        l += ec.l[4] * al;
        r += ec.r[3] * dm;
      }
      pert_ppm1_sel(l, r);
      if (ec.on) {
        bl = l;
        br = r;
      }
    } else {
      // the left edge face of this cell is the right one of its left neighbour
      const int role = edge_role(i, is, ie, npx, edge);
      const T d1 = from_left(dxl), d3 = from_right(dxl), d4 = from_right(d3);
      const T efr = edge_face(qm1, q, qp1, qp2, d1, dxl, d3, d4);
      if constexpr (SIDE == 0)
        edge_coef_sel(qm1, q, qp1, dmm1, dm, dmp1, al, alp1, role, from_left(efr), efr, bl, br);
      else
        edge_coef_side<SIDE>(qm1, q, qp1, dmm1, dm, dmp1, al, alp1, role, from_left(efr), efr, bl, br);
    }
  }
  const T sum = bl + br;
  return face_flux_s(c, qm1, from_left(br), from_left(sum), q, bl, sum);
}

// Sliding y-direction PPM over a column: fed one cell per row, yields the
// flux at face r-2 once rows r-4..r are in.  Histories are indexed by the
// row modulo 6 so that an unrolled loop keeps them in fixed registers.
template <typename T>
struct YPipe {
  T w[6];   // q of rows, slot row%6
  T dm[6];  // dm of cells
  T al[6];  // face values
  T bl[6], br[6], bs[6];  // bs = bl + br
};

// One step at input row r; U is the slot of row r.  role/dd: the cube-edge
// role of cell r-2 and the spacings of cells r-4..r (uniform over the wave).
// TW: the row's index t when it is one of a block's first six input rows (the
// window then holds rows 0..t only, and each stage runs from the first row at
// which its inputs exist: dm from t = 2, face values from 3, cell
// coefficients from 4, the flux from 5 -- the first face the block needs);
// 6 in the steady state.  The flux is 0 before t = 5.
template <typename T, int U, int TW = 6, bool EXT = false>
__device__ __forceinline__ T ypipe_step(YPipe<T> &p, T qnew, const T *wx, T c, int iord, int role, T d0, T d1,
                                        T d2, T d3, T d4) {
  constexpr int s0 = U % 6, s1 = (U + 5) % 6, s2 = (U + 4) % 6, s3 = (U + 3) % 6,
                s4 = (U + 2) % 6;  // rows r, r-1, r-2, r-3, r-4
  // the window's cells: held in the pipe, or (EXT) given as rows r-1 .. r-4
  const T w0 = qnew;
  if constexpr (!EXT) p.w[s0] = qnew;
  const T w1 = EXT ? wx[0] : p.w[s1], w2 = EXT ? wx[1] : p.w[s2], w3 = EXT ? wx[2] : p.w[s3],
          w4 = EXT ? wx[3] : p.w[s4];
  if constexpr (TW >= 2) p.dm[s1] = dm3(w2, w1, w0);
  if constexpr (TW >= 3) p.al[s1] = al2(w2, w1, p.dm[s2], p.dm[s1]);
  if constexpr (TW >= 4) {
    cell_coef(w4, w3, w2, w1, w0, p.dm[s3], p.dm[s2], p.dm[s1], p.al[s2], p.al[s1], iord, p.bl[s2], p.br[s2]);
    if (role != 0) {  // uniform over the wave
      const T efl = edge_face(w4, w3, w2, w1, d0, d1, d2, d3);
      const T efr = edge_face(w3, w2, w1, w0, d1, d2, d3, d4);
      edge_coef_sel(w3, w2, w1, p.dm[s3], p.dm[s2], p.dm[s1], p.al[s2], p.al[s1], role, efl, efr, p.bl[s2],
                    p.br[s2]);
    }
    p.bs[s2] = p.bl[s2] + p.br[s2];
  }
  if constexpr (TW >= 5)
    return face_flux_s(c, w3, p.br[s3], p.bs[s3], w2, p.bl[s2], p.bs[s2]);
  else
    return T(0);
}

}  // namespace fv3
