// mapz_column.h -- one-column vertical remap on the device: the volume-
// conserving cubic-spline interface solve with the kord 9/10 constraints
// (upstream fv_mapz.F90 cs_profile / scalar_profile / cs_limiters), the
// monotone source-layer search with exact parabola integrals (map1_ppm,
// map1_q2, mapn_tracer), and fillz.  Expressions keep the upstream evaluation
// order so that T = double reproduces an FP64 CPU evaluation bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fv3_device.h"

namespace fv3 {

constexpr int REMAP_MAX_NQ = 64;

template <typename T>
struct RemapArgs {
  int mode;  // 0: Lagrangian_to_Eulerian tracer subset; 1: single field, given pe1/pe2
  int km, kn, nq, ntiles, ni, nj;
  long long ncol;
  int is, js, isd, jsd, nid;
  long long s2;
  const double *ak, *bk;
  double ptop;
  double *delp;
  T *q, *pt;
  int kord[REMAP_MAX_NQ];
  int kord_tm;
  T t_min;
  int fill;
  bool mapn;
  // mode 1
  const double *pe1_in, *pe2_in, *dp2_in;
  const T *q1_in, *qs_in;
  T *q2_out;
  long long q1_stride, q2_stride;  // level strides
  int iv, kord1, variant;          // variant 0 map1_ppm, 1 map_scalar, 2 map1_q2
  bool use_qmin;
  T qmin;
  // scratch, [level][column]
  double *pe1, *pe2, *dp1, *dp2;
  T *a4, *qi, *gam, *dm;
  uint8_t *flags;
  int *err;
};

template <typename T>
struct CV {  // 1-based view of one column with level stride st
  T *p;
  long long st;
  __device__ __forceinline__ T &operator()(int k) const { return p[(long long)(k - 1) * st]; }
};

// cs_limiters on one cell
template <typename T>
__device__ __forceinline__ void cs_lim1(T &a1, T &a2, T &a3, T &a4, bool extm, int iv) {
  if (iv == 0) {
    if (a1 <= T(0)) {
      a2 = a1;
      a3 = a1;
      a4 = T(0);
    } else if (fabs(a3 - a2) < -a4) {
      // a1 + (a3-a2)^2 / (4 a4) + a4 / 12 < 0; a4 < 0 here, so the FP32 form
      // tests the same sign multiplied through by a4, without a division
      bool dip;
      if constexpr (sizeof(T) == 4)
        dip = fmaf(a1, a4, T(0.25) * ((a3 - a2) * (a3 - a2)) + a4 * a4 * K<T>::r12) > T(0);
      else
        dip = (a1 + T(0.25) * ((a3 - a2) * (a3 - a2)) / a4 + a4 * K<T>::r12) < T(0);
      if (dip) {
        if (a1 < a3 && a1 < a2) {
          a3 = a1;
          a2 = a1;
          a4 = T(0);
        } else if (a3 > a2) {
          a4 = T(3) * (a2 - a1);
          a3 = a2 - a4;
        } else {
          a4 = T(3) * (a3 - a1);
          a2 = a3 - a4;
        }
      }
    }
    return;
  }
  bool flat = iv == 1 ? (a1 - a2) * (a1 - a3) >= T(0) : extm;
  if (flat) {
    a2 = a1;
    a3 = a1;
    a4 = T(0);
  } else {
    T da1 = a3 - a2;
    T da2 = da1 * da1;
    T a6da = a4 * da1;
    if (a6da < -da2) {
      a4 = T(3) * (a2 - a1);
      a3 = a2 - a4;
    } else if (a6da > da2) {
      a4 = T(3) * (a3 - a1);
      a2 = a3 - a4;
    }
  }
}

template <typename T>
__device__ __forceinline__ void huynh2(T a1, T &a2, T &a3, T gk, T gk1, T gk2, T gkm1) {
  T pmp_1 = a1 - T(2) * gk1;
  T lac_1 = pmp_1 + T(1.5) * gk2;
  a2 = mn2(mx2(a2, mn3(a1, pmp_1, lac_1)), mx3(a1, pmp_1, lac_1));
  T pmp_2 = a1 + T(2) * gk;
  T lac_2 = pmp_2 - T(1.5) * gkm1;
  a3 = mn2(mx2(a3, mn3(a1, pmp_2, lac_2)), mx3(a1, pmp_2, lac_2));
}

// cs_profile (use_qmin false) / scalar_profile (use_qmin true) on one column.
// A1..A4: a4(1..4, k); DP: layer thickness (FP64); Q, G: km+1 work levels;
// E: 3*km flags (extm, ext5, ext6).
template <typename T>
__device__ void cs_profile_col(const CV<T> &A1, const CV<T> &A2, const CV<T> &A3, const CV<T> &A4,
                               const CV<double> &DP, T qs, int km, int iv, int kord,
                               bool use_qmin, T qmin, const CV<T> &Q, const CV<T> &G,
                               const CV<uint8_t> &EXTM, const CV<uint8_t> &EXT5,
                               const CV<uint8_t> &EXT6) {
  const int akord = kord < 0 ? -kord : kord;
  if (iv == -2) {
    G(2) = T(0.5);
    Q(1) = T(1.5) * A1(1);
    for (int k = 2; k <= km - 1; k++) {
      T grat = T(DP(k - 1) / DP(k));
      T bet = T(2) + grat + grat - G(k);
      Q(k) = (T(3) * (A1(k - 1) + A1(k)) - Q(k - 1)) / bet;
      G(k + 1) = grat / bet;
    }
    T grat = T(DP(km - 1) / DP(km));
    Q(km) = (T(3) * (A1(km - 1) + A1(km)) - grat * qs - Q(km - 1)) / (T(2) + grat + grat - G(km));
    Q(km + 1) = qs;
    for (int k = km - 1; k >= 1; k--) Q(k) = Q(k) - G(k + 1) * Q(k + 1);
  } else {
    T grat = T(DP(2) / DP(1));
    T bet = grat * (grat + T(0.5));
    Q(1) = ((grat + grat) * (grat + T(1)) * A1(1) + A1(2)) / bet;
    G(1) = (T(1) + grat * (grat + T(1.5))) / bet;
    T d4 = T(0);
    for (int k = 2; k <= km; k++) {
      d4 = T(DP(k - 1) / DP(k));
      T bt = T(2) + d4 + d4 - G(k - 1);
      Q(k) = (T(3) * (A1(k - 1) + d4 * A1(k)) - Q(k - 1)) / bt;
      G(k) = d4 / bt;
    }
    T a_bot = T(1) + d4 * (d4 + T(1.5));
    Q(km + 1) = (T(2) * d4 * (d4 + T(1)) * A1(km) + A1(km - 1) - a_bot * Q(km)) /
                (d4 * (d4 + T(0.5)) - a_bot * G(km));
    for (int k = km; k >= 1; k--) Q(k) = Q(k) - G(k) * Q(k + 1);
  }
  if (akord > 16) {
    for (int k = 1; k <= km; k++) {
      A2(k) = Q(k);
      A3(k) = Q(k + 1);
      A4(k) = T(3) * (T(2) * A1(k) - (A2(k) + A3(k)));
    }
    return;
  }
  Q(2) = mn2(Q(2), mx2(A1(1), A1(2)));
  Q(2) = mx2(Q(2), mn2(A1(1), A1(2)));
  for (int k = 2; k <= km; k++) G(k) = A1(k) - A1(k - 1);
  for (int k = 3; k <= km - 1; k++) {
    if (G(k - 1) * G(k + 1) > T(0)) {
      Q(k) = mn2(Q(k), mx2(A1(k - 1), A1(k)));
      Q(k) = mx2(Q(k), mn2(A1(k - 1), A1(k)));
    } else if (G(k - 1) > T(0)) {
      Q(k) = mx2(Q(k), mn2(A1(k - 1), A1(k)));
    } else {
      Q(k) = mn2(Q(k), mx2(A1(k - 1), A1(k)));
      if (iv == 0) Q(k) = mx2(T(0), Q(k));
    }
  }
  Q(km) = mn2(Q(km), mx2(A1(km - 1), A1(km)));
  Q(km) = mx2(Q(km), mn2(A1(km - 1), A1(km)));
  for (int k = 1; k <= km; k++) {
    A2(k) = Q(k);
    A3(k) = Q(k + 1);
  }
  for (int k = 1; k <= km; k++) {
    if (k == 1 || k == km)
      EXTM(k) = (A2(k) - A1(k)) * (A3(k) - A1(k)) > T(0);
    else
      EXTM(k) = G(k) * G(k + 1) < T(0);
    if (akord > 9) {
      T x0 = T(2) * A1(k) - (A2(k) + A3(k));
      T x1 = fabs(A2(k) - A3(k));
      A4(k) = T(3) * x0;
      EXT5(k) = fabs(x0) > x1;
      EXT6(k) = fabs(A4(k)) > x1;
    }
  }
  if (iv == 0) {
    A2(1) = mx2(T(0), A2(1));
  } else if (iv == -1) {
    if (A2(1) * A1(1) <= T(0)) A2(1) = T(0);
  } else if (iv == 2) {
    A2(1) = A1(1);
    A3(1) = A1(1);
    A4(1) = T(0);
  }
  if (iv != 2) {
    A4(1) = T(3) * (T(2) * A1(1) - (A2(1) + A3(1)));
    cs_lim1(A1(1), A2(1), A3(1), A4(1), EXTM(1) != 0, 1);
  }
  A4(2) = T(3) * (T(2) * A1(2) - (A2(2) + A3(2)));
  cs_lim1(A1(2), A2(2), A3(2), A4(2), EXTM(2) != 0, 2);
  for (int k = 3; k <= km - 2; k++) {
    T a1 = A1(k), a2 = A2(k), a3 = A3(k), a4 = A4(k);
    if (akord < 9) {
      huynh2(a1, a2, a3, G(k), G(k + 1), G(k + 2), G(k - 1));
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    } else if (akord == 9) {
      const bool em = EXTM(k) != 0;
      if ((em && EXTM(k - 1)) || (em && EXTM(k + 1)) || (use_qmin && em && a1 < qmin)) {
        a2 = a1;
        a3 = a1;
        a4 = T(0);
      } else {
        a4 = T(6) * a1 - T(3) * (a2 + a3);
        if (fabs(a4) > fabs(a2 - a3)) {
          huynh2(a1, a2, a3, G(k), G(k + 1), G(k + 2), G(k - 1));
          a4 = T(6) * a1 - T(3) * (a2 + a3);
        }
      }
    } else {  // kord 10
      if (EXT5(k)) {
        if (EXT5(k - 1) || EXT5(k + 1) || (use_qmin && a1 < qmin)) {
          a2 = a1;
          a3 = a1;
        } else if (EXT6(k - 1) || EXT6(k + 1)) {
          huynh2(a1, a2, a3, G(k), G(k + 1), G(k + 2), G(k - 1));
        }
      } else if (EXT6(k)) {
        if (EXT5(k - 1) || EXT5(k + 1)) huynh2(a1, a2, a3, G(k), G(k + 1), G(k + 2), G(k - 1));
      }
      a4 = T(3) * (T(2) * a1 - (a2 + a3));
    }
    if (iv == 0) cs_lim1(a1, a2, a3, a4, EXTM(k) != 0, 0);
    A2(k) = a2;
    A3(k) = a3;
    A4(k) = a4;
  }
  if (iv == 0) {
    A3(km) = mx2(T(0), A3(km));
  } else if (iv == -1) {
    if (A3(km) * A1(km) <= T(0)) A3(km) = T(0);
  }
  for (int k = km - 1; k <= km; k++) {
    A4(k) = T(3) * (T(2) * A1(k) - (A2(k) + A3(k)));
    cs_lim1(A1(k), A2(k), A3(k), A4(k), EXTM(k) != 0, k == km - 1 ? 2 : 1);
  }
}

// Source-layer search and parabola integrals onto kn target layers.
// mapn: mapn_tracer's factored form; otherwise map1_ppm / map1_q2's.
// Division by den(k) = dp2 when given, else pe2(k+1) - pe2(k).
template <typename T, typename OUT>
__device__ void map_col(const CV<double> &PE1, const CV<double> &PE2, const CV<double> &D1,
                        const CV<T> &A1, const CV<T> &A2, const CV<T> &A3, const CV<T> &A4,
                        int km, int kn, bool mapn, const CV<double> *DEN, OUT out) {
  const double r3 = 1.0 / 3.0, r23 = 2.0 / 3.0;
  int k0 = 1;
  for (int k = 1; k <= kn; k++) {
    double qsum = 0.0;
    bool single = false;
    for (int l = k0; l <= km; l++) {
      if (PE2(k) >= PE1(l) && PE2(k) <= PE1(l + 1)) {
        const double pl = (PE2(k) - PE1(l)) / D1(l);
        if (PE2(k + 1) <= PE1(l + 1)) {
          const double pr = (PE2(k + 1) - PE1(l)) / D1(l);
          T v;
          if (mapn) {
            double fac1 = pr + pl;
            double fac2 = r3 * (pr * fac1 + pl * pl);
            fac1 = 0.5 * fac1;
            v = A2(l) + (A4(l) + A3(l) - A2(l)) * T(fac1) - A4(l) * T(fac2);
          } else {
            v = A2(l) + T(0.5) * (A4(l) + A3(l) - A2(l)) * T(pr + pl) -
                A4(l) * T(r3) * T(pr * (pr + pl) + pl * pl);
          }
          out(k, v);
          k0 = l;
          single = true;
          break;
        }
        if (mapn) {
          double dp = PE1(l + 1) - PE2(k);
          double fac1 = 1.0 + pl;
          double fac2 = r3 * (1.0 + pl * fac1);
          fac1 = 0.5 * fac1;
          qsum = dp * (double)(A2(l) + (A4(l) + A3(l) - A2(l)) * T(fac1) - A4(l) * T(fac2));
        } else {
          qsum = (PE1(l + 1) - PE2(k)) *
                 (double)(A2(l) + T(0.5) * (A4(l) + A3(l) - A2(l)) * T(1.0 + pl) -
                          A4(l) * T(r3 * (1.0 + pl * (1.0 + pl))));
        }
        for (int m = l + 1; m <= km; m++) {
          if (PE2(k + 1) > PE1(m + 1)) {
            qsum = qsum + D1(m) * (double)A1(m);
          } else {
            const double dp = PE2(k + 1) - PE1(m);
            const double esl = dp / D1(m);
            if (mapn) {
              const double f1 = 0.5 * esl, f2 = 1.0 - r23 * esl;
              qsum = qsum + dp * (double)(A2(m) + T(f1) * (A3(m) - A2(m) + A4(m) * T(f2)));
            } else {
              qsum = qsum + dp * (double)(A2(m) + T(0.5 * esl) * (A3(m) - A2(m) + A4(m) * T(1.0 - r23 * esl)));
            }
            k0 = m;
            break;
          }
        }
        break;
      }
    }
    if (!single) {
      const double den = DEN ? (*DEN)(k) : PE2(k + 1) - PE2(k);
      out(k, T(qsum / den));
    }
  }
}

// fillz on one column of one tracer
template <typename T>
__device__ void fillz_col(const CV<T> &Q, const CV<double> &DP, const CV<T> &DM, int km) {
  if (Q(1) < T(0)) {
    Q(2) = T((double)Q(2) + (double)Q(1) * DP(1) / DP(2));
    Q(1) = T(0);
  }
  bool zfix = false;
  for (int k = 2; k <= km - 1; k++) {
    if (Q(k) < T(0)) {
      zfix = true;
      if (Q(k - 1) > T(0)) {
        double dq = mn2((double)Q(k - 1) * DP(k - 1), -(double)Q(k) * DP(k));
        Q(k - 1) = T((double)Q(k - 1) - dq / DP(k - 1));
        Q(k) = T((double)Q(k) + dq / DP(k));
      }
      if (Q(k) < T(0) && Q(k + 1) > T(0)) {
        double dq = mn2((double)Q(k + 1) * DP(k + 1), -(double)Q(k) * DP(k));
        Q(k + 1) = T((double)Q(k + 1) - dq / DP(k + 1));
        Q(k) = T((double)Q(k) + dq / DP(k));
      }
    }
  }
  if (Q(km) < T(0) && Q(km - 1) > T(0)) {
    zfix = true;
    double qup = (double)Q(km - 1) * DP(km - 1);
    double qly = -(double)Q(km) * DP(km);
    double dup = mn2(qly, qup);
    Q(km - 1) = T((double)Q(km - 1) - dup / DP(km - 1));
    Q(km) = T((double)Q(km) + dup / DP(km));
  }
  if (zfix) {
    double sum0 = 0.0;
    for (int k = 2; k <= km; k++) {
      DM(k) = T((double)Q(k) * DP(k));
      sum0 = sum0 + (double)DM(k);
    }
    if (sum0 > 0.0) {
      double sum1 = 0.0;
      for (int k = 2; k <= km; k++) sum1 = sum1 + mx2(0.0, (double)DM(k));
      double fac = sum0 / sum1;
      for (int k = 2; k <= km; k++) Q(k) = T(mx2(0.0, fac * (double)DM(k) / DP(k)));
    }
  }
}

template <typename T>
__device__ void remap_column(const RemapArgs<T> &a, long long col) {
  const long long nc = a.ncol;
  const int km = a.km;
  CV<double> PE1{a.pe1 + col, nc}, PE2{a.pe2 + col, nc}, D1{a.dp1 + col, nc}, D2{a.dp2 + col, nc};
  CV<T> A1{a.a4 + col, 4 * nc}, A2{a.a4 + nc + col, 4 * nc}, A3{a.a4 + 2 * nc + col, 4 * nc},
      A4{a.a4 + 3 * nc + col, 4 * nc};
  CV<T> Q{a.qi + col, nc}, G{a.gam + col, nc}, DM{a.dm + col, nc};
  CV<uint8_t> EXTM{a.flags + col, 3 * nc}, EXT5{a.flags + nc + col, 3 * nc},
      EXT6{a.flags + 2 * nc + col, 3 * nc};
  if (a.mode == 1) {
    CV<double> P1{const_cast<double *>(a.pe1_in) + col, nc}, P2{const_cast<double *>(a.pe2_in) + col, nc};
    for (int k = 1; k <= km; k++) {
      D1(k) = P1(k + 1) - P1(k);
      A1(k) = a.q1_in[(long long)(k - 1) * a.q1_stride + col];
    }
    const T qs = a.qs_in ? a.qs_in[col] : T(0);
    cs_profile_col<T>(A1, A2, A3, A4, D1, qs, km, a.iv, a.kord1, a.use_qmin, a.qmin, Q, G, EXTM, EXT5, EXT6);
    T *o = a.q2_out + col;
    const long long os = a.q2_stride;
    auto out = [=](int k, T v) { o[(long long)(k - 1) * os] = v; };
    if (a.variant == 2 || a.variant == 3) {
      CV<double> DN{const_cast<double *>(a.dp2_in) + col, nc};
      map_col<T>(P1, P2, D1, A1, A2, A3, A4, km, a.kn, a.variant == 3, &DN, out);
      if (a.variant == 3 && a.fill) fillz_col<T>(CV<T>{o, os}, DN, DM, km);
    } else {
      map_col<T>(P1, P2, D1, A1, A2, A3, A4, km, a.kn, false, (const CV<double> *)nullptr, out);
    }
    return;
  }
  const long long per_tile = (long long)a.ni * a.nj;
  const int tile = (int)(col / per_tile);
  const long long r = col % per_tile;
  const int i = a.is + (int)(r % a.ni), j = a.js + (int)(r / a.ni);
  const long long cell = (long long)(j - a.jsd) * a.nid + (i - a.isd);
  double *dl = a.delp + (long long)tile * km * a.s2 + cell;
  PE1(1) = a.ptop;
  for (int k = 1; k <= km; k++) PE1(k + 1) = PE1(k) + dl[(long long)(k - 1) * a.s2];
  PE2(1) = a.ptop;
  PE2(km + 1) = PE1(km + 1);
  for (int k = 2; k <= km; k++) PE2(k) = a.ak[k - 1] + a.bk[k - 1] * PE1(km + 1);
  for (int k = 1; k <= km; k++) {
    D2(k) = PE2(k + 1) - PE2(k);
    D1(k) = PE1(k + 1) - PE1(k);
  }
  for (int iq = 0; iq < a.nq; iq++) {
    T *qc = a.q + (long long)tile * km * a.nq * a.s2 + (long long)iq * km * a.s2 + cell;
    for (int k = 1; k <= km; k++) A1(k) = qc[(long long)(k - 1) * a.s2];
    cs_profile_col<T>(A1, A2, A3, A4, D1, T(0), km, 0, a.kord[iq], true, T(0), Q, G, EXTM, EXT5, EXT6);
    const long long s2 = a.s2;
    auto out = [=](int k, T v) { qc[(long long)(k - 1) * s2] = v; };
    map_col<T>(PE1, PE2, D1, A1, A2, A3, A4, km, km, a.mapn, &D2, out);
    if (a.fill) fillz_col<T>(CV<T>{qc, a.s2}, D2, DM, km);
  }
  if (a.pt) {
    T *pc = a.pt + (long long)tile * km * a.s2 + cell;
    for (int k = 1; k <= km; k++) A1(k) = pc[(long long)(k - 1) * a.s2];
    cs_profile_col<T>(A1, A2, A3, A4, D1, T(0), km, 1, a.kord_tm, true, a.t_min, Q, G, EXTM, EXT5, EXT6);
    const long long s2 = a.s2;
    auto out = [=](int k, T v) { pc[(long long)(k - 1) * s2] = v; };
    map_col<T>(PE1, PE2, D1, A1, A2, A3, A4, km, km, false, (const CV<double> *)nullptr, out);
  }
  for (int k = 1; k <= km; k++) dl[(long long)(k - 1) * a.s2] = D2(k);
}

struct Ctx;
// launch remap_column over a.ncol columns with scratch from the context
template <typename T>
int remap_launch(Ctx &c, RemapArgs<T> &a);

}  // namespace fv3
