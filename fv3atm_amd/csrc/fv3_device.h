// fv3_device.h -- device building blocks of the MI355X tracer transport path:
// the per-face PPM flux of xppm/yppm (upstream tp_core.F90), pert_ppm, and the
// copy_corners index maps.  Every expression keeps the evaluation order of
// the upstream Fortran statement so that the FP64 instantiation, built with
// contraction off, reproduces a CPU FP64 evaluation bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fv3 {

constexpr int NG = 3;

template <typename T>
struct K {
  static constexpr T r3 = T(1.0 / 3.0);
  static constexpr T near_zero = T(1.0e-25);
  static constexpr T ppm_fac = T(1.5);
  static constexpr T s11 = T(11.0 / 14.0);
  static constexpr T s14 = T(4.0 / 7.0);
  static constexpr T s15 = T(3.0 / 14.0);
  static constexpr T r12 = T(1.0 / 12.0);
};

template <typename T> __host__ __device__ __forceinline__ T mx2(T a, T b) { return a > b ? a : b; }
template <typename T> __host__ __device__ __forceinline__ T mn2(T a, T b) { return a < b ? a : b; }
template <typename T> __host__ __device__ __forceinline__ T mx3(T a, T b, T c) { return mx2(mx2(a, b), c); }
template <typename T> __host__ __device__ __forceinline__ T mn3(T a, T b, T c) { return mn2(mn2(a, b), c); }
template <typename T> __host__ __device__ __forceinline__ T mx4(T a, T b, T c, T d) { return mx2(mx3(a, b, c), d); }
template <typename T> __host__ __device__ __forceinline__ T mn4(T a, T b, T c, T d) { return mn2(mn3(a, b, c), d); }
// Fortran SIGN(a, b)
template <typename T> __host__ __device__ __forceinline__ T fsign(T a, T b) { return copysign(fabs(a), b); }

// Static description of one sweep line: which cells get the cube-edge
// treatment of xppm/yppm.
struct LineGeom {
  int is, ie, npx;  // compute range and corner-point count along the line
  bool edge;        // !bounded_domain && grid_type < 3
};

// Lin's monotone slope dm(m)
template <typename T, typename Q>
__device__ __forceinline__ T ppm_dm(const Q &q, int m) {
  T qm = q(m - 1), q0 = q(m), qp = q(m + 1);
  T xt = T(0.25) * (qp - qm);
  return fsign(mn3(fabs(xt), mx3(qm, q0, qp) - q0, q0 - mn3(qm, q0, qp)), xt);
}

// face value al(m) (between cells m-1 and m)
template <typename T, typename Q>
__device__ __forceinline__ T ppm_al(const Q &q, int m, T dm_lo, T dm_hi) {
  return T(0.5) * (q(m - 1) + q(m)) + K<T>::r3 * (dm_lo - dm_hi);
}

// pert_ppm on one cell
template <typename T>
__device__ __forceinline__ void pert_ppm1(T a0, T &al, T &ar, int iv) {
  if (iv == 0) {
    if (a0 <= T(0)) {
      al = T(0);
      ar = T(0);
    } else {
      T a4 = T(-3) * (ar + al);
      T da1 = ar - al;
      if (fabs(da1) < -a4) {
        T fmin = a0 + T(0.25) / a4 * (da1 * da1) + a4 * K<T>::r12;
        if (fmin < T(0)) {
          if (ar > T(0) && al > T(0)) {
            ar = T(0);
            al = T(0);
          } else if (da1 > T(0)) {
            ar = T(-2) * al;
          } else {
            al = T(-2) * ar;
          }
        }
      }
    }
  } else {
    if (al * ar < T(0)) {
      T da1 = al - ar;
      T da2 = da1 * da1;
      T a6da = T(3) * (al + ar) * da1;
      if (a6da < -da2)
        ar = T(-2) * al;
      else if (a6da > da2)
        al = T(-2) * ar;
    } else {
      al = T(0);
      ar = T(0);
    }
  }
}

// Cube-edge interface value with the local 4-cell clamp (xppm/yppm edge code):
// face between cells a and a+1, with cells a-1..a+2 and spacings d(a-1..a+2).
template <typename T, typename Q, typename D>
__device__ __forceinline__ T ppm_edge_face(const Q &q, const D &d, int a) {
  T qm = q(a - 1), q0 = q(a), q1 = q(a + 1), q2 = q(a + 2);
  T dm = d(a - 1), d0 = d(a), d1 = d(a + 1), d2 = d(a + 2);
  T xt = T(0.5) * (((T(2) * d0 + dm) * q0 - d0 * qm) / (dm + d0) +
                   ((T(2) * d1 + d2) * q1 - d1 * q2) / (d1 + d2));
  xt = mx2(xt, mn4(qm, q0, q1, q2));
  xt = mn2(xt, mx4(qm, q0, q1, q2));
  return xt;
}

// (bl, br) of cell m for the monotone orders iord 7..13.
template <typename T, typename Q, typename D>
__device__ __forceinline__ void ppm_cell(const Q &q, const D &d, int m, const LineGeom &g,
                                         int iord, T &bl, T &br) {
  const int npx = g.npx;
  if (g.edge && g.is == 1 && m <= 2) {
    if (m == 0) {
      bl = K<T>::s14 * ppm_dm<T>(q, -1) + K<T>::s11 * (q(-1) - q(0));
      br = ppm_edge_face<T>(q, d, 0) - q(0);
    } else if (m == 1) {
      bl = ppm_edge_face<T>(q, d, 0) - q(1);
      br = K<T>::s15 * q(1) + K<T>::s11 * q(2) - K<T>::s14 * ppm_dm<T>(q, 2) - q(1);
    } else {
      bl = K<T>::s15 * q(1) + K<T>::s11 * q(2) - K<T>::s14 * ppm_dm<T>(q, 2) - q(2);
      br = ppm_al<T>(q, 3, ppm_dm<T>(q, 2), ppm_dm<T>(q, 3)) - q(2);
    }
    pert_ppm1(q(m), bl, br, 1);
    return;
  }
  if (g.edge && g.ie + 1 == npx && m >= npx - 2) {
    if (m == npx - 2) {
      bl = ppm_al<T>(q, npx - 2, ppm_dm<T>(q, npx - 3), ppm_dm<T>(q, npx - 2)) - q(npx - 2);
      br = K<T>::s15 * q(npx - 1) + K<T>::s11 * q(npx - 2) + K<T>::s14 * ppm_dm<T>(q, npx - 2) -
           q(npx - 2);
    } else if (m == npx - 1) {
      bl = K<T>::s15 * q(npx - 1) + K<T>::s11 * q(npx - 2) + K<T>::s14 * ppm_dm<T>(q, npx - 2) -
           q(npx - 1);
      br = ppm_edge_face<T>(q, d, npx - 1) - q(npx - 1);
    } else {
      bl = ppm_edge_face<T>(q, d, npx - 1) - q(npx);
      br = K<T>::s11 * (q(npx + 1) - q(npx)) - K<T>::s14 * ppm_dm<T>(q, npx + 1);
    }
    pert_ppm1(q(m), bl, br, 1);
    return;
  }
  const T dmm = ppm_dm<T>(q, m - 1), dm0 = ppm_dm<T>(q, m), dmp = ppm_dm<T>(q, m + 1);
  const T q0 = q(m);
  const T all = ppm_al<T>(q, m, dmm, dm0), alr = ppm_al<T>(q, m + 1, dm0, dmp);
  if (iord == 8 || iord == 11) {
    T xt = iord == 8 ? T(2) * dm0 : K<T>::ppm_fac * dm0;
    bl = -fsign(mn2(fabs(xt), fabs(all - q0)), xt);
    br = fsign(mn2(fabs(xt), fabs(alr - q0)), xt);
  } else {
    bl = all - q0;
    br = alr - q0;
    if (fabs(dmm) + fabs(dm0) + fabs(dmp) < K<T>::near_zero) {
      bl = T(0);
      br = T(0);
    } else if (fabs(T(3) * (bl + br)) > fabs(bl - br)) {
      T dqm2 = T(2) * (q(m - 1) - q(m - 2));
      T dqm1 = T(2) * (q0 - q(m - 1));
      T dq0 = T(2) * (q(m + 1) - q0);
      T dqp1 = T(2) * (q(m + 2) - q(m + 1));
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

// Upwind PPM face flux at face i (between cells i-1 and i) for Courant c.
template <typename T, typename Q, typename D>
__device__ __forceinline__ T ppm_flux(const Q &q, const D &d, int i, T c, const LineGeom &g,
                                      int iord) {
  T bl, br;
  if (c > T(0)) {
    ppm_cell<T>(q, d, i - 1, g, iord, bl, br);
    return q(i - 1) + (T(1) - c) * (br - c * (bl + br));
  } else {
    ppm_cell<T>(q, d, i, g, iord, bl, br);
    return q(i) + (T(1) + c) * (bl + c * (bl + br));
  }
}

// copy_corners source of corner-block cell (i, j); returns false outside the
// corner blocks of the corners this domain owns.
__device__ __host__ __forceinline__ bool corner_src(int i, int j, int npx, int npy, int dir,
                                                     unsigned corners, int &si, int &sj) {
  const bool sw = corners & 1, se = corners & 2, ne = corners & 4, nw = corners & 8;
  const bool ilo = i <= 0 && i >= 1 - NG, ihi = i >= npx && i <= npx + NG - 1;
  const bool jlo = j <= 0 && j >= 1 - NG, jhi = j >= npy && j <= npy + NG - 1;
  if (dir == 1) {
    if (sw && ilo && jlo) { si = j; sj = 1 - i; return true; }
    if (se && ihi && jlo) { si = npy - j; sj = i - npx + 1; return true; }
    if (ne && ihi && jhi) { si = j; sj = 2 * npx - 1 - i; return true; }
    if (nw && ilo && jhi) { si = npy - j; sj = i - 1 + npx; return true; }
  } else {
    if (sw && ilo && jlo) { si = 1 - j; sj = i; return true; }
    if (se && ihi && jlo) { si = npy + j - 1; sj = npx - i; return true; }
    if (ne && ihi && jhi) { si = 2 * npy - 1 - j; sj = i; return true; }
    if (nw && ilo && jhi) { si = j + 1 - npx; sj = npy - i; return true; }
  }
  return false;
}

}  // namespace fv3
