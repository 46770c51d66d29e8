/*
 * fv3_oracle_mapz.c -- FP64 CPU restatement of FV3's vertical remap of
 * tracers and pt (upstream GFDL_atmos_cubed_sphere model/fv_mapz.F90:
 * Lagrangian_to_Eulerian, mapn_tracer, map1_ppm, map1_q2, map_scalar,
 * cs_profile, scalar_profile, cs_limiters, fillz; the submodule is empty in
 * the reference, .gitmodules:1-4, and its importers are
 * moving_nest/fv_moving_nest.F90:85 and ccpp/data/CCPP_typedefs.F90:1714).
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP product.
 *
 * Floating-point expressions keep the Fortran statement's evaluation order.
 */
#include "fv3_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

static const double R3 = 1.0 / 3.0;
static const double R23 = 2.0 / 3.0;
static const double R12 = 1.0 / 12.0;

static inline double mx2(double a, double b) { return a > b ? a : b; }
static inline double mn2(double a, double b) { return a < b ? a : b; }
static inline double mx3(double a, double b, double c) { return mx2(mx2(a, b), c); }
static inline double mn3(double a, double b, double c) { return mn2(mn2(a, b), c); }

/* a4(4, i1:i2, km): element (m, i, k) */
#define A4(m, i, k) a4[((size_t)((k) - 1) * im + ((i) - i1)) * 4 + ((m) - 1)]
#define DP(i, k) delp[(size_t)((k) - 1) * im + ((i) - i1)]

/* cs_limiters (fv_mapz.F90) on im cells of one layer; a4 is (4, im). */
void fv3o_cs_limiters(int im, const int *extm, double *a4, int iv) {
#define L4(m, i) a4[(size_t)(i) * 4 + ((m) - 1)]
  if (iv == 0) {
    for (int i = 0; i < im; i++) {
      if (L4(1, i) <= 0.0) {
        L4(2, i) = L4(1, i);
        L4(3, i) = L4(1, i);
        L4(4, i) = 0.0;
      } else if (fabs(L4(3, i) - L4(2, i)) < -L4(4, i)) {
        if ((L4(1, i) + 0.25 * ((L4(3, i) - L4(2, i)) * (L4(3, i) - L4(2, i))) / L4(4, i) +
             L4(4, i) * R12) < 0.0) {
          if (L4(1, i) < L4(3, i) && L4(1, i) < L4(2, i)) {
            L4(3, i) = L4(1, i);
            L4(2, i) = L4(1, i);
            L4(4, i) = 0.0;
          } else if (L4(3, i) > L4(2, i)) {
            L4(4, i) = 3.0 * (L4(2, i) - L4(1, i));
            L4(3, i) = L4(2, i) - L4(4, i);
          } else {
            L4(4, i) = 3.0 * (L4(3, i) - L4(1, i));
            L4(2, i) = L4(3, i) - L4(4, i);
          }
        }
      }
    }
  } else if (iv == 1) {
    for (int i = 0; i < im; i++) {
      if ((L4(1, i) - L4(2, i)) * (L4(1, i) - L4(3, i)) >= 0.0) {
        L4(2, i) = L4(1, i);
        L4(3, i) = L4(1, i);
        L4(4, i) = 0.0;
      } else {
        double da1 = L4(3, i) - L4(2, i);
        double da2 = da1 * da1;
        double a6da = L4(4, i) * da1;
        if (a6da < -da2) {
          L4(4, i) = 3.0 * (L4(2, i) - L4(1, i));
          L4(3, i) = L4(2, i) - L4(4, i);
        } else if (a6da > da2) {
          L4(4, i) = 3.0 * (L4(3, i) - L4(1, i));
          L4(2, i) = L4(3, i) - L4(4, i);
        }
      }
    }
  } else {
    for (int i = 0; i < im; i++) {
      if (extm[i]) {
        L4(2, i) = L4(1, i);
        L4(3, i) = L4(1, i);
        L4(4, i) = 0.0;
      } else {
        double da1 = L4(3, i) - L4(2, i);
        double da2 = da1 * da1;
        double a6da = L4(4, i) * da1;
        if (a6da < -da2) {
          L4(4, i) = 3.0 * (L4(2, i) - L4(1, i));
          L4(3, i) = L4(2, i) - L4(4, i);
        } else if (a6da > da2) {
          L4(4, i) = 3.0 * (L4(3, i) - L4(1, i));
          L4(2, i) = L4(3, i) - L4(4, i);
        }
      }
    }
  }
#undef L4
}

/* Huynh's 2nd constraint on both edges of layer k (the shared block of the
 * kord < 9, 9 and 10 branches). */
static inline void huynh_edges(double *a4k, const double *gam, int k, int km, int i, int im) {
  (void)km;
#define G(kk) gam[(size_t)((kk) - 1) * im + i]
  double pmp_1 = a4k[0] - 2.0 * G(k + 1);
  double lac_1 = pmp_1 + 1.5 * G(k + 2);
  a4k[1] = mn2(mx2(a4k[1], mn3(a4k[0], pmp_1, lac_1)), mx3(a4k[0], pmp_1, lac_1));
  double pmp_2 = a4k[0] + 2.0 * G(k);
  double lac_2 = pmp_2 - 1.5 * G(k - 1);
  a4k[2] = mn2(mx2(a4k[2], mn3(a4k[0], pmp_2, lac_2)), mx3(a4k[0], pmp_2, lac_2));
#undef G
}

/* cs_profile / scalar_profile (fv_mapz.F90).  use_qmin selects the
 * scalar_profile floor tests (kord 9 and 10 flatten an extremum whose mean is
 * below qmin). */
static int cs_profile_impl(const double *qs, double *a4, const double *delp, int km, int i1,
                           int i2, int iv, int kord, int use_qmin, double qmin) {
  const int im = i2 - i1 + 1;
  const int akord = abs(kord);
  if (km < 4) return 1;
  if (!(akord <= 10 || akord > 16)) return 1;
  double *q = (double *)malloc(sizeof(double) * (size_t)im * (km + 1));
  double *gam = (double *)malloc(sizeof(double) * (size_t)im * km);
  double *d4 = (double *)malloc(sizeof(double) * (size_t)im);
  int *extm = (int *)malloc(sizeof(int) * (size_t)im * km);
  int *ext5 = (int *)malloc(sizeof(int) * (size_t)im * km);
  int *ext6 = (int *)malloc(sizeof(int) * (size_t)im * km);
  int *exk = (int *)malloc(sizeof(int) * (size_t)im);
#define QI(i, k) q[(size_t)((k) - 1) * im + ((i) - i1)]
#define GAM(i, k) gam[(size_t)((k) - 1) * im + ((i) - i1)]
#define EXTM(i, k) extm[(size_t)((k) - 1) * im + ((i) - i1)]
#define EXT5(i, k) ext5[(size_t)((k) - 1) * im + ((i) - i1)]
#define EXT6(i, k) ext6[(size_t)((k) - 1) * im + ((i) - i1)]

  if (iv == -2) {
    for (int i = i1; i <= i2; i++) {
      GAM(i, 2) = 0.5;
      QI(i, 1) = 1.5 * A4(1, i, 1);
    }
    for (int k = 2; k <= km - 1; k++)
      for (int i = i1; i <= i2; i++) {
        double grat = DP(i, k - 1) / DP(i, k);
        double bet = 2.0 + grat + grat - GAM(i, k);
        QI(i, k) = (3.0 * (A4(1, i, k - 1) + A4(1, i, k)) - QI(i, k - 1)) / bet;
        GAM(i, k + 1) = grat / bet;
      }
    for (int i = i1; i <= i2; i++) {
      double grat = DP(i, km - 1) / DP(i, km);
      QI(i, km) = (3.0 * (A4(1, i, km - 1) + A4(1, i, km)) - grat * qs[i - i1] - QI(i, km - 1)) /
                  (2.0 + grat + grat - GAM(i, km));
      QI(i, km + 1) = qs[i - i1];
    }
    for (int k = km - 1; k >= 1; k--)
      for (int i = i1; i <= i2; i++) QI(i, k) = QI(i, k) - GAM(i, k + 1) * QI(i, k + 1);
  } else {
    for (int i = i1; i <= i2; i++) {
      double grat = DP(i, 2) / DP(i, 1);
      double bet = grat * (grat + 0.5);
      QI(i, 1) = ((grat + grat) * (grat + 1.0) * A4(1, i, 1) + A4(1, i, 2)) / bet;
      GAM(i, 1) = (1.0 + grat * (grat + 1.5)) / bet;
    }
    for (int k = 2; k <= km; k++)
      for (int i = i1; i <= i2; i++) {
        d4[i - i1] = DP(i, k - 1) / DP(i, k);
        double bet = 2.0 + d4[i - i1] + d4[i - i1] - GAM(i, k - 1);
        QI(i, k) = (3.0 * (A4(1, i, k - 1) + d4[i - i1] * A4(1, i, k)) - QI(i, k - 1)) / bet;
        GAM(i, k) = d4[i - i1] / bet;
      }
    for (int i = i1; i <= i2; i++) {
      double dd = d4[i - i1];
      double a_bot = 1.0 + dd * (dd + 1.5);
      QI(i, km + 1) = (2.0 * dd * (dd + 1.0) * A4(1, i, km) + A4(1, i, km - 1) - a_bot * QI(i, km)) /
                      (dd * (dd + 0.5) - a_bot * GAM(i, km));
    }
    for (int k = km; k >= 1; k--)
      for (int i = i1; i <= i2; i++) QI(i, k) = QI(i, k) - GAM(i, k) * QI(i, k + 1);
  }

  if (akord > 16) { /* perfectly linear scheme */
    for (int k = 1; k <= km; k++)
      for (int i = i1; i <= i2; i++) {
        A4(2, i, k) = QI(i, k);
        A4(3, i, k) = QI(i, k + 1);
        A4(4, i, k) = 3.0 * (2.0 * A4(1, i, k) - (A4(2, i, k) + A4(3, i, k)));
      }
    goto done;
  }

  /* large-scale constraints */
  for (int i = i1; i <= i2; i++) {
    QI(i, 2) = mn2(QI(i, 2), mx2(A4(1, i, 1), A4(1, i, 2)));
    QI(i, 2) = mx2(QI(i, 2), mn2(A4(1, i, 1), A4(1, i, 2)));
  }
  for (int k = 2; k <= km; k++)
    for (int i = i1; i <= i2; i++) GAM(i, k) = A4(1, i, k) - A4(1, i, k - 1);
  for (int k = 3; k <= km - 1; k++)
    for (int i = i1; i <= i2; i++) {
      if (GAM(i, k - 1) * GAM(i, k + 1) > 0.0) {
        QI(i, k) = mn2(QI(i, k), mx2(A4(1, i, k - 1), A4(1, i, k)));
        QI(i, k) = mx2(QI(i, k), mn2(A4(1, i, k - 1), A4(1, i, k)));
      } else if (GAM(i, k - 1) > 0.0) {
        QI(i, k) = mx2(QI(i, k), mn2(A4(1, i, k - 1), A4(1, i, k)));
      } else {
        QI(i, k) = mn2(QI(i, k), mx2(A4(1, i, k - 1), A4(1, i, k)));
        if (iv == 0) QI(i, k) = mx2(0.0, QI(i, k));
      }
    }
  for (int i = i1; i <= i2; i++) {
    QI(i, km) = mn2(QI(i, km), mx2(A4(1, i, km - 1), A4(1, i, km)));
    QI(i, km) = mx2(QI(i, km), mn2(A4(1, i, km - 1), A4(1, i, km)));
  }
  for (int k = 1; k <= km; k++)
    for (int i = i1; i <= i2; i++) {
      A4(2, i, k) = QI(i, k);
      A4(3, i, k) = QI(i, k + 1);
    }
  for (int k = 1; k <= km; k++) {
    for (int i = i1; i <= i2; i++) {
      if (k == 1 || k == km)
        EXTM(i, k) = (A4(2, i, k) - A4(1, i, k)) * (A4(3, i, k) - A4(1, i, k)) > 0.0;
      else
        EXTM(i, k) = GAM(i, k) * GAM(i, k + 1) < 0.0;
    }
    if (akord > 9)
      for (int i = i1; i <= i2; i++) {
        double x0 = 2.0 * A4(1, i, k) - (A4(2, i, k) + A4(3, i, k));
        double x1 = fabs(A4(2, i, k) - A4(3, i, k));
        A4(4, i, k) = 3.0 * x0;
        EXT5(i, k) = fabs(x0) > x1;
        EXT6(i, k) = fabs(A4(4, i, k)) > x1;
      }
  }

  /* top two layers: monotone mapping */
  if (iv == 0) {
    for (int i = i1; i <= i2; i++) A4(2, i, 1) = mx2(0.0, A4(2, i, 1));
  } else if (iv == -1) {
    for (int i = i1; i <= i2; i++)
      if (A4(2, i, 1) * A4(1, i, 1) <= 0.0) A4(2, i, 1) = 0.0;
  } else if (iv == 2) {
    for (int i = i1; i <= i2; i++) {
      A4(2, i, 1) = A4(1, i, 1);
      A4(3, i, 1) = A4(1, i, 1);
      A4(4, i, 1) = 0.0;
    }
  }
  if (iv != 2) {
    for (int i = i1; i <= i2; i++)
      A4(4, i, 1) = 3.0 * (2.0 * A4(1, i, 1) - (A4(2, i, 1) + A4(3, i, 1)));
    for (int i = i1; i <= i2; i++) exk[i - i1] = EXTM(i, 1);
    fv3o_cs_limiters(im, exk, &A4(1, i1, 1), 1);
  }
  for (int i = i1; i <= i2; i++)
    A4(4, i, 2) = 3.0 * (2.0 * A4(1, i, 2) - (A4(2, i, 2) + A4(3, i, 2)));
  for (int i = i1; i <= i2; i++) exk[i - i1] = EXTM(i, 2);
  fv3o_cs_limiters(im, exk, &A4(1, i1, 2), 2);

  /* interior: Huynh's 2nd constraint */
  for (int k = 3; k <= km - 2; k++) {
    if (akord < 9) {
      for (int i = i1; i <= i2; i++) {
        huynh_edges(&A4(1, i, k), gam, k, km, i - i1, im);
        A4(4, i, k) = 3.0 * (2.0 * A4(1, i, k) - (A4(2, i, k) + A4(3, i, k)));
      }
    } else if (akord == 9) {
      for (int i = i1; i <= i2; i++) {
        if ((EXTM(i, k) && EXTM(i, k - 1)) || (EXTM(i, k) && EXTM(i, k + 1)) ||
            (use_qmin && EXTM(i, k) && A4(1, i, k) < qmin)) {
          A4(2, i, k) = A4(1, i, k);
          A4(3, i, k) = A4(1, i, k);
          A4(4, i, k) = 0.0;
        } else {
          A4(4, i, k) = 6.0 * A4(1, i, k) - 3.0 * (A4(2, i, k) + A4(3, i, k));
          if (fabs(A4(4, i, k)) > fabs(A4(2, i, k) - A4(3, i, k))) {
            huynh_edges(&A4(1, i, k), gam, k, km, i - i1, im);
            A4(4, i, k) = 6.0 * A4(1, i, k) - 3.0 * (A4(2, i, k) + A4(3, i, k));
          }
        }
      }
    } else if (akord == 10) {
      for (int i = i1; i <= i2; i++) {
        if (EXT5(i, k)) {
          if (EXT5(i, k - 1) || EXT5(i, k + 1) || (use_qmin && A4(1, i, k) < qmin)) {
            A4(2, i, k) = A4(1, i, k);
            A4(3, i, k) = A4(1, i, k);
          } else if (EXT6(i, k - 1) || EXT6(i, k + 1)) {
            huynh_edges(&A4(1, i, k), gam, k, km, i - i1, im);
          }
        } else if (EXT6(i, k)) {
          if (EXT5(i, k - 1) || EXT5(i, k + 1)) huynh_edges(&A4(1, i, k), gam, k, km, i - i1, im);
        }
      }
      for (int i = i1; i <= i2; i++)
        A4(4, i, k) = 3.0 * (2.0 * A4(1, i, k) - (A4(2, i, k) + A4(3, i, k)));
    }
    if (iv == 0) {
      for (int i = i1; i <= i2; i++) exk[i - i1] = EXTM(i, k);
      fv3o_cs_limiters(im, exk, &A4(1, i1, k), 0);
    }
  }

  /* bottom two layers */
  if (iv == 0) {
    for (int i = i1; i <= i2; i++) A4(3, i, km) = mx2(0.0, A4(3, i, km));
  } else if (iv == -1) {
    for (int i = i1; i <= i2; i++)
      if (A4(3, i, km) * A4(1, i, km) <= 0.0) A4(3, i, km) = 0.0;
  }
  for (int k = km - 1; k <= km; k++) {
    for (int i = i1; i <= i2; i++)
      A4(4, i, k) = 3.0 * (2.0 * A4(1, i, k) - (A4(2, i, k) + A4(3, i, k)));
    for (int i = i1; i <= i2; i++) exk[i - i1] = EXTM(i, k);
    fv3o_cs_limiters(im, exk, &A4(1, i1, k), k == km - 1 ? 2 : 1);
  }

done:
  free(q);
  free(gam);
  free(d4);
  free(extm);
  free(ext5);
  free(ext6);
  free(exk);
#undef QI
#undef GAM
#undef EXTM
#undef EXT5
#undef EXT6
  return 0;
}

int fv3o_cs_profile(const double *qs, double *a4, const double *delp, int km, int i1, int i2,
                    int iv, int kord) {
  return cs_profile_impl(qs, a4, delp, km, i1, i2, iv, kord, 0, 0.0);
}

int fv3o_scalar_profile(const double *qs, double *a4, const double *delp, int km, int i1,
                        int i2, int iv, int kord, double qmin) {
  return cs_profile_impl(qs, a4, delp, km, i1, i2, iv, kord, 1, qmin);
}

/* Mapping step shared by map1_ppm / map_scalar / map1_q2 (fv_mapz.F90):
 * monotone search for the source layer holding each target top edge, then
 * the exact integral of the parabola over the overlap.  Output q2 column i
 * layer k is written to out[(k-1)*ostride + (i-i1)] divided by
 * den(i,k) (pe2 difference, or dp2 when dp2 != NULL). */
static void map_columns(int km, const double *pe1, const double *a4, const double *dp1, int kn,
                        const double *pe2, int i1, int i2, double *out, size_t ostride,
                        const double *dp2) {
  const int im = i2 - i1 + 1;
#define PE1(i, k) pe1[(size_t)((k) - 1) * im + ((i) - i1)]
#define PE2(i, k) pe2[(size_t)((k) - 1) * im + ((i) - i1)]
#define D1(i, k) dp1[(size_t)((k) - 1) * im + ((i) - i1)]
  for (int i = i1; i <= i2; i++) {
    int k0 = 1;
    for (int k = 1; k <= kn; k++) {
      double qsum = 0.0;
      int single = 0;
      for (int l = k0; l <= km; l++) {
        if (PE2(i, k) >= PE1(i, l) && PE2(i, k) <= PE1(i, l + 1)) {
          double pl = (PE2(i, k) - PE1(i, l)) / D1(i, l);
          if (PE2(i, k + 1) <= PE1(i, l + 1)) {
            double pr = (PE2(i, k + 1) - PE1(i, l)) / D1(i, l);
            out[(size_t)(k - 1) * ostride + (i - i1)] =
                A4(2, i, l) + 0.5 * (A4(4, i, l) + A4(3, i, l) - A4(2, i, l)) * (pr + pl) -
                A4(4, i, l) * R3 * (pr * (pr + pl) + pl * pl);
            k0 = l;
            single = 1;
            break;
          }
          qsum = (PE1(i, l + 1) - PE2(i, k)) *
                 (A4(2, i, l) + 0.5 * (A4(4, i, l) + A4(3, i, l) - A4(2, i, l)) * (1.0 + pl) -
                  A4(4, i, l) * (R3 * (1.0 + pl * (1.0 + pl))));
          for (int m = l + 1; m <= km; m++) {
            if (PE2(i, k + 1) > PE1(i, m + 1)) {
              qsum = qsum + D1(i, m) * A4(1, i, m);
            } else {
              double dp = PE2(i, k + 1) - PE1(i, m);
              double esl = dp / D1(i, m);
              qsum = qsum + dp * (A4(2, i, m) + 0.5 * esl *
                                      (A4(3, i, m) - A4(2, i, m) + A4(4, i, m) * (1.0 - R23 * esl)));
              k0 = m;
              break;
            }
          }
          break;
        }
      }
      if (!single) {
        double den = dp2 ? dp2[(size_t)(k - 1) * im + (i - i1)] : PE2(i, k + 1) - PE2(i, k);
        out[(size_t)(k - 1) * ostride + (i - i1)] = qsum / den;
      }
    }
  }
#undef PE1
#undef PE2
#undef D1
}

static int map_generic(int km, const double *pe1, const double *q1, const double *qs, int kn,
                       const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                       int jbeg, int jend, int iv, int kord, int scalar, double q_min) {
  const int im = i2 - i1 + 1;
  const size_t ni = (size_t)(iend - ibeg + 1), plane = ni * (jend - jbeg + 1);
  double *dp1 = (double *)malloc(sizeof(double) * (size_t)im * km);
  double *a4 = (double *)malloc(sizeof(double) * 4 * (size_t)im * km);
  double *out = (double *)malloc(sizeof(double) * (size_t)im * kn);
  for (int k = 1; k <= km; k++)
    for (int i = i1; i <= i2; i++) {
      dp1[(size_t)(k - 1) * im + (i - i1)] =
          pe1[(size_t)k * im + (i - i1)] - pe1[(size_t)(k - 1) * im + (i - i1)];
      A4(1, i, k) = q1[(size_t)(k - 1) * plane + (size_t)(j - jbeg) * ni + (i - ibeg)];
    }
  int rc;
  if (kord > 7)
    rc = scalar ? fv3o_scalar_profile(qs, a4, dp1, km, i1, i2, iv, kord, q_min)
                : fv3o_cs_profile(qs, a4, dp1, km, i1, i2, iv, kord);
  else
    rc = 1; /* ppm_profile (kord <= 7) is not part of this path */
  if (!rc) {
    map_columns(km, pe1, a4, dp1, kn, pe2, i1, i2, out, (size_t)im, NULL);
    for (int k = 1; k <= kn; k++)
      for (int i = i1; i <= i2; i++)
        q2[(size_t)(k - 1) * plane + (size_t)(j - jbeg) * ni + (i - ibeg)] =
            out[(size_t)(k - 1) * im + (i - i1)];
  }
  free(dp1);
  free(a4);
  free(out);
  return rc;
}

int fv3o_map1_ppm(int km, const double *pe1, const double *q1, const double *qs, int kn,
                  const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                  int jbeg, int jend, int iv, int kord) {
  return map_generic(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord,
                     0, 0.0);
}

int fv3o_map_scalar(int km, const double *pe1, const double *q1, const double *qs, int kn,
                    const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                    int jbeg, int jend, int iv, int kord, double q_min) {
  return map_generic(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord,
                     1, q_min);
}

/* map1_q2: one tracer of a row into q2(i1:i2, kn), dividing by dp2. */
int fv3o_map1_q2(int km, const double *pe1, const double *q1, int kn, const double *pe2,
                 double *q2, const double *dp2, int i1, int i2, int iv, int kord, int j,
                 int ibeg, int iend, int jbeg, int jend, double q_min) {
  const int im = i2 - i1 + 1;
  const size_t ni = (size_t)(iend - ibeg + 1), plane = ni * (jend - jbeg + 1);
  double *dp1 = (double *)malloc(sizeof(double) * (size_t)im * km);
  double *a4 = (double *)malloc(sizeof(double) * 4 * (size_t)im * km);
  for (int k = 1; k <= km; k++)
    for (int i = i1; i <= i2; i++) {
      dp1[(size_t)(k - 1) * im + (i - i1)] =
          pe1[(size_t)k * im + (i - i1)] - pe1[(size_t)(k - 1) * im + (i - i1)];
      A4(1, i, k) = q1[(size_t)(k - 1) * plane + (size_t)(j - jbeg) * ni + (i - ibeg)];
    }
  int rc = kord > 7 ? fv3o_scalar_profile(NULL, a4, dp1, km, i1, i2, iv, kord, q_min) : 1;
  if (!rc) map_columns(km, pe1, a4, dp1, kn, pe2, i1, i2, q2, (size_t)im, dp2);
  free(dp1);
  free(a4);
  return rc;
}

/* mapn_tracer: all nq tracers of a row share the layer search; fac1/fac2
 * factor the parabola integral once per overlap. */
int fv3o_mapn_tracer(int nq, int km, const double *pe1, const double *pe2, double *q1,
                     const double *dp2, const int *kord, int j, int i1, int i2, int isd,
                     int ied, int jsd, int jed, double q_min, int fill) {
  const int im = i2 - i1 + 1;
  const size_t ni = (size_t)(ied - isd + 1), plane = ni * (jed - jsd + 1);
  const size_t colsz = (size_t)im * km;
  double *dp1 = (double *)malloc(sizeof(double) * colsz);
  double *q4 = (double *)malloc(sizeof(double) * 4 * colsz * nq);
  double *q2 = (double *)malloc(sizeof(double) * colsz * nq);
  double *qsum = (double *)malloc(sizeof(double) * nq);
  int rc = 0;
#define PE1(i, k) pe1[(size_t)((k) - 1) * im + ((i) - i1)]
#define PE2(i, k) pe2[(size_t)((k) - 1) * im + ((i) - i1)]
#define D1(i, k) dp1[(size_t)((k) - 1) * im + ((i) - i1)]
#define Q4(m, i, k, iq) q4[(((size_t)((iq) - 1) * km + ((k) - 1)) * im + ((i) - i1)) * 4 + ((m) - 1)]
#define Q2(i, k, iq) q2[((size_t)((iq) - 1) * km + ((k) - 1)) * im + ((i) - i1)]
  for (int k = 1; k <= km; k++)
    for (int i = i1; i <= i2; i++) D1(i, k) = PE1(i, k + 1) - PE1(i, k);
  for (int iq = 1; iq <= nq; iq++) {
    for (int k = 1; k <= km; k++)
      for (int i = i1; i <= i2; i++)
        Q4(1, i, k, iq) =
            q1[((size_t)(iq - 1) * km + (k - 1)) * plane + (size_t)(j - jsd) * ni + (i - isd)];
    if (kord[iq - 1] <= 7) {
      rc = 1;
      goto out;
    }
    rc |= fv3o_scalar_profile(NULL, &Q4(1, i1, 1, iq), dp1, km, i1, i2, 0, kord[iq - 1], q_min);
  }
  for (int i = i1; i <= i2; i++) {
    int k0 = 1;
    for (int k = 1; k <= km; k++) {
      int single = 0;
      for (int l = k0; l <= km; l++) {
        if (PE2(i, k) >= PE1(i, l) && PE2(i, k) <= PE1(i, l + 1)) {
          double pl = (PE2(i, k) - PE1(i, l)) / D1(i, l);
          if (PE2(i, k + 1) <= PE1(i, l + 1)) {
            double pr = (PE2(i, k + 1) - PE1(i, l)) / D1(i, l);
            double fac1 = pr + pl;
            double fac2 = R3 * (pr * fac1 + pl * pl);
            fac1 = 0.5 * fac1;
            for (int iq = 1; iq <= nq; iq++)
              Q2(i, k, iq) = Q4(2, i, l, iq) +
                             (Q4(4, i, l, iq) + Q4(3, i, l, iq) - Q4(2, i, l, iq)) * fac1 -
                             Q4(4, i, l, iq) * fac2;
            k0 = l;
            single = 1;
            break;
          }
          double dp = PE1(i, l + 1) - PE2(i, k);
          double fac1 = 1.0 + pl;
          double fac2 = R3 * (1.0 + pl * fac1);
          fac1 = 0.5 * fac1;
          for (int iq = 1; iq <= nq; iq++)
            qsum[iq - 1] = dp * (Q4(2, i, l, iq) +
                                 (Q4(4, i, l, iq) + Q4(3, i, l, iq) - Q4(2, i, l, iq)) * fac1 -
                                 Q4(4, i, l, iq) * fac2);
          for (int m = l + 1; m <= km; m++) {
            if (PE2(i, k + 1) > PE1(i, m + 1)) {
              for (int iq = 1; iq <= nq; iq++) qsum[iq - 1] = qsum[iq - 1] + D1(i, m) * Q4(1, i, m, iq);
            } else {
              double dpm = PE2(i, k + 1) - PE1(i, m);
              double esl = dpm / D1(i, m);
              double f1 = 0.5 * esl;
              double f2 = 1.0 - R23 * esl;
              for (int iq = 1; iq <= nq; iq++)
                qsum[iq - 1] = qsum[iq - 1] +
                               dpm * (Q4(2, i, m, iq) +
                                      f1 * (Q4(3, i, m, iq) - Q4(2, i, m, iq) + Q4(4, i, m, iq) * f2));
              k0 = m;
              break;
            }
          }
          break;
        }
      }
      if (!single)
        for (int iq = 1; iq <= nq; iq++)
          Q2(i, k, iq) = qsum[iq - 1] / dp2[(size_t)(k - 1) * im + (i - i1)];
    }
  }
  if (fill) fv3o_fillz(im, km, nq, q2, dp2);
  for (int iq = 1; iq <= nq; iq++)
    for (int k = 1; k <= km; k++)
      for (int i = i1; i <= i2; i++)
        q1[((size_t)(iq - 1) * km + (k - 1)) * plane + (size_t)(j - jsd) * ni + (i - isd)] =
            Q2(i, k, iq);
out:
#undef PE1
#undef PE2
#undef D1
#undef Q4
#undef Q2
  free(dp1);
  free(q4);
  free(q2);
  free(qsum);
  return rc;
}

/* fillz (fv_mapz.F90): column-conserving removal of negative values;
 * q(im, km, nq), dp(im, km). */
void fv3o_fillz(int im, int km, int nq, double *q, const double *dp) {
  int *zfix = (int *)malloc(sizeof(int) * (size_t)im);
  double *dm = (double *)malloc(sizeof(double) * (size_t)(km + 1));
#define QF(i, k, n) q[((size_t)(n) * km + ((k) - 1)) * im + (i)]
#define DPF(i, k) dp[(size_t)((k) - 1) * im + (i)]
  for (int n = 0; n < nq; n++) {
    for (int i = 0; i < im; i++)
      if (QF(i, 1, n) < 0.0) {
        QF(i, 2, n) = QF(i, 2, n) + QF(i, 1, n) * DPF(i, 1) / DPF(i, 2);
        QF(i, 1, n) = 0.0;
      }
    for (int i = 0; i < im; i++) zfix[i] = 0;
    for (int k = 2; k <= km - 1; k++)
      for (int i = 0; i < im; i++)
        if (QF(i, k, n) < 0.0) {
          zfix[i] = 1;
          if (QF(i, k - 1, n) > 0.0) {
            double dq = mn2(QF(i, k - 1, n) * DPF(i, k - 1), -QF(i, k, n) * DPF(i, k));
            QF(i, k - 1, n) = QF(i, k - 1, n) - dq / DPF(i, k - 1);
            QF(i, k, n) = QF(i, k, n) + dq / DPF(i, k);
          }
          if (QF(i, k, n) < 0.0 && QF(i, k + 1, n) > 0.0) {
            double dq = mn2(QF(i, k + 1, n) * DPF(i, k + 1), -QF(i, k, n) * DPF(i, k));
            QF(i, k + 1, n) = QF(i, k + 1, n) - dq / DPF(i, k + 1);
            QF(i, k, n) = QF(i, k, n) + dq / DPF(i, k);
          }
        }
    for (int i = 0; i < im; i++)
      if (QF(i, km, n) < 0.0 && QF(i, km - 1, n) > 0.0) {
        zfix[i] = 1;
        double qup = QF(i, km - 1, n) * DPF(i, km - 1);
        double qly = -QF(i, km, n) * DPF(i, km);
        double dup = mn2(qly, qup);
        QF(i, km - 1, n) = QF(i, km - 1, n) - dup / DPF(i, km - 1);
        QF(i, km, n) = QF(i, km, n) + dup / DPF(i, km);
      }
    for (int i = 0; i < im; i++) {
      if (!zfix[i]) continue;
      double sum0 = 0.0;
      for (int k = 2; k <= km; k++) {
        dm[k] = QF(i, k, n) * DPF(i, k);
        sum0 = sum0 + dm[k];
      }
      if (sum0 > 0.0) {
        double sum1 = 0.0;
        for (int k = 2; k <= km; k++) sum1 = sum1 + mx2(0.0, dm[k]);
        double fac = sum0 / sum1;
        for (int k = 2; k <= km; k++) QF(i, k, n) = mx2(0.0, fac * dm[k] / DPF(i, k));
      }
    }
  }
#undef QF
#undef DPF
  free(zfix);
  free(dm);
}

/* Tracer/pt subset of Lagrangian_to_Eulerian (fv_mapz.F90), all rows. */
int fv3o_lagrangian_to_eulerian_tracers(int km, int nq, const fv3o_bounds *bd,
                                        const double *ak, const double *bk, double ptop,
                                        double *delp, double *q, double *pt,
                                        const int *kord_tr, int kord_tm, double t_min,
                                        int fill) {
  const int is = bd->is, ie = bd->ie, im = ie - is + 1;
  const int isd = bd->isd, ied = bd->ied, jsd = bd->jsd, jed = bd->jed;
  const size_t ni = (size_t)(ied - isd + 1), plane = ni * (jed - jsd + 1);
  int rc = 0;
  /* rows are independent (upstream's OpenMP loop over j) */
#pragma omp parallel reduction(| : rc)
  {
    double *pe1 = (double *)malloc(sizeof(double) * (size_t)im * (km + 1));
    double *pe2 = (double *)malloc(sizeof(double) * (size_t)im * (km + 1));
    double *dp2 = (double *)malloc(sizeof(double) * (size_t)im * km);
    double *q2 = (double *)malloc(sizeof(double) * (size_t)im * km);
#define PE1(i, k) pe1[(size_t)((k) - 1) * im + ((i) - is)]
#define PE2(i, k) pe2[(size_t)((k) - 1) * im + ((i) - is)]
#pragma omp for schedule(dynamic, 1)
    for (int j = bd->js; j <= bd->je; j++) {
      for (int i = is; i <= ie; i++) PE1(i, 1) = ptop;
      for (int k = 1; k <= km; k++)
        for (int i = is; i <= ie; i++)
          PE1(i, k + 1) = PE1(i, k) + delp[(size_t)(k - 1) * plane + (size_t)(j - jsd) * ni + (i - isd)];
      for (int i = is; i <= ie; i++) {
        PE2(i, 1) = ptop;
        PE2(i, km + 1) = PE1(i, km + 1);
      }
      for (int k = 2; k <= km; k++)
        for (int i = is; i <= ie; i++) PE2(i, k) = ak[k - 1] + bk[k - 1] * PE1(i, km + 1);
      for (int k = 1; k <= km; k++)
        for (int i = is; i <= ie; i++)
          dp2[(size_t)(k - 1) * im + (i - is)] = PE2(i, k + 1) - PE2(i, k);

      if (nq > 5) {
        rc |= fv3o_mapn_tracer(nq, km, pe1, pe2, q, dp2, kord_tr, j, is, ie, isd, ied, jsd, jed,
                               0.0, fill);
      } else {
        for (int iq = 1; iq <= nq; iq++) {
          double *qq = q + (size_t)(iq - 1) * km * plane;
          rc |= fv3o_map1_q2(km, pe1, qq, km, pe2, q2, dp2, is, ie, 0, kord_tr[iq - 1], j, isd,
                             ied, jsd, jed, 0.0);
          if (fill) fv3o_fillz(im, km, 1, q2, dp2);
          for (int k = 1; k <= km; k++)
            for (int i = is; i <= ie; i++)
              qq[(size_t)(k - 1) * plane + (size_t)(j - jsd) * ni + (i - isd)] =
                  q2[(size_t)(k - 1) * im + (i - is)];
        }
      }
      if (pt != NULL)
        rc |= fv3o_map_scalar(km, pe1, pt, NULL, km, pe2, pt, is, ie, j, isd, ied, jsd, jed, 1,
                              kord_tm, t_min);
      for (int k = 1; k <= km; k++)
        for (int i = is; i <= ie; i++)
          delp[(size_t)(k - 1) * plane + (size_t)(j - jsd) * ni + (i - isd)] =
              dp2[(size_t)(k - 1) * im + (i - is)];
    }
#undef PE1
#undef PE2
    free(pe1);
    free(pe2);
    free(dp2);
    free(q2);
  }
  return rc;
}
