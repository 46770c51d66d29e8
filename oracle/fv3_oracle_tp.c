/*
 * fv3_oracle_tp.c -- FP64 CPU restatement of FV3 tracer transport
 * (upstream GFDL_atmos_cubed_sphere model/tp_core.F90 and
 * model/fv_tracer2d.F90; that submodule is empty in the reference,
 * .gitmodules:1-4).  TEST INFRASTRUCTURE ONLY -- the checker for the HIP
 * product, never linked into it.
 *
 * Every floating-point expression keeps the left-to-right evaluation order of
 * the Fortran statement it restates, so that the HIP FP64 form (compiled with
 * contraction off, like this file) can reproduce it to the last bit.
 */
#include "fv3_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#define NG 3

/* tp_core constants (upstream module parameters). */
static const double R3 = 1.0 / 3.0;
static const double NEAR_ZERO = 1.0e-25;
static const double PPM_FAC = 1.5;
static const double S11 = 11.0 / 14.0, S14 = 4.0 / 7.0, S15 = 3.0 / 14.0;

static inline double mx2(double a, double b) { return a > b ? a : b; }
static inline double mn2(double a, double b) { return a < b ? a : b; }
static inline double mx3(double a, double b, double c) { return mx2(mx2(a, b), c); }
static inline double mn3(double a, double b, double c) { return mn2(mn2(a, b), c); }
static inline double mx4(double a, double b, double c, double d) { return mx2(mx3(a, b, c), d); }
static inline double mn4(double a, double b, double c, double d) { return mn2(mn3(a, b, c), d); }
/* Fortran SIGN(a, b) */
static inline double fsign(double a, double b) { return copysign(fabs(a), b); }

/* 2-D array element with Fortran lower bounds (il, jl) and leading extent ni */
#define E2(p, i, j, il, jl, ni) (p)[(ptrdiff_t)((j) - (jl)) * (ni) + ((i) - (il))]

/* ------------------------------------------------------------------------ */
/* pert_ppm (tp_core.F90): perturbation-form PPM constraint on (a0, al, ar). */
void fv3o_pert_ppm(int im, const double *a0, double *al, double *ar, int iv) {
  const double r12 = 1.0 / 12.0;
  if (iv == 0) {
    for (int i = 0; i < im; i++) {
      if (a0[i] <= 0.0) {
        al[i] = 0.0;
        ar[i] = 0.0;
      } else {
        double a4 = -3.0 * (ar[i] + al[i]);
        double da1 = ar[i] - al[i];
        if (fabs(da1) < -a4) {
          double fmin = a0[i] + 0.25 / a4 * (da1 * da1) + a4 * r12;
          if (fmin < 0.0) {
            if (ar[i] > 0.0 && al[i] > 0.0) {
              ar[i] = 0.0;
              al[i] = 0.0;
            } else if (da1 > 0.0) {
              ar[i] = -2.0 * al[i];
            } else {
              al[i] = -2.0 * ar[i];
            }
          }
        }
      }
    }
  } else {
    for (int i = 0; i < im; i++) {
      if (al[i] * ar[i] < 0.0) {
        double da1 = al[i] - ar[i];
        double da2 = da1 * da1;
        double a6da = 3.0 * (al[i] + ar[i]) * da1;
        if (a6da < -da2) {
          ar[i] = -2.0 * al[i];
        } else if (a6da > da2) {
          al[i] = -2.0 * ar[i];
        }
      } else {
        al[i] = 0.0;
        ar[i] = 0.0;
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* One line of xppm / yppm (tp_core.F90), monotone family iord 7..13.
 * q1, dxa1 are indexed [isd..ied]; c1 and flux1 [is..ie+1]. */
static int ppm_line(double *flux1, const double *q1, const double *c1, const double *dxa1,
                    int iord, int is, int ie, int isd, int ied, int npx, int bounded_domain,
                    int grid_type) {
  if (iord < 7 || iord > 13) return 1;
  const int lo = isd - 1, n = ied - isd + 3;
  double *buf = (double *)malloc(sizeof(double) * 5 * (size_t)n);
  double *dm = buf, *al = buf + n, *bl = buf + 2 * n, *br = buf + 3 * n, *dq = buf + 4 * n;
  for (int t = 0; t < 5 * n; t++) buf[t] = NAN; /* poison: unset entries must not be read */
#define Q(i) q1[(i) - isd]
#define DXA(i) dxa1[(i) - isd]
#define L(a, i) a[(i) - lo]
  const int edge = (!bounded_domain && grid_type < 3);
  int is1, ie1;
  if (edge) {
    is1 = is - 1 > 3 ? is - 1 : 3;
    ie1 = npx - 3 < ie + 1 ? npx - 3 : ie + 1;
  } else {
    is1 = is - 1;
    ie1 = ie + 1;
  }

  for (int i = is - 2; i <= ie + 2; i++) {
    double xt = 0.25 * (Q(i + 1) - Q(i - 1));
    L(dm, i) = fsign(mn3(fabs(xt), mx3(Q(i - 1), Q(i), Q(i + 1)) - Q(i),
                         Q(i) - mn3(Q(i - 1), Q(i), Q(i + 1))),
                     xt);
  }
  for (int i = is1; i <= ie1 + 1; i++)
    L(al, i) = 0.5 * (Q(i - 1) + Q(i)) + R3 * (L(dm, i - 1) - L(dm, i));

  if (iord == 8) {
    for (int i = is1; i <= ie1; i++) {
      double xt = 2.0 * L(dm, i);
      L(bl, i) = -fsign(mn2(fabs(xt), fabs(L(al, i) - Q(i))), xt);
      L(br, i) = fsign(mn2(fabs(xt), fabs(L(al, i + 1) - Q(i))), xt);
    }
  } else if (iord == 11) {
    for (int i = is1; i <= ie1; i++) {
      double xt = PPM_FAC * L(dm, i);
      L(bl, i) = -fsign(mn2(fabs(xt), fabs(L(al, i) - Q(i))), xt);
      L(br, i) = fsign(mn2(fabs(xt), fabs(L(al, i + 1) - Q(i))), xt);
    }
  } else {
    for (int i = is1 - 2; i <= ie1 + 1; i++) L(dq, i) = 2.0 * (Q(i + 1) - Q(i));
    for (int i = is1; i <= ie1; i++) {
      L(bl, i) = L(al, i) - Q(i);
      L(br, i) = L(al, i + 1) - Q(i);
      if (fabs(L(dm, i - 1)) + fabs(L(dm, i)) + fabs(L(dm, i + 1)) < NEAR_ZERO) {
        L(bl, i) = 0.0;
        L(br, i) = 0.0;
      } else if (fabs(3.0 * (L(bl, i) + L(br, i))) > fabs(L(bl, i) - L(br, i))) {
        double pmp_2 = L(dq, i - 1);
        double lac_2 = pmp_2 - 0.75 * L(dq, i - 2);
        L(br, i) = mn2(mx3(0.0, pmp_2, lac_2), mx2(L(br, i), mn3(0.0, pmp_2, lac_2)));
        double pmp_1 = -L(dq, i);
        double lac_1 = pmp_1 + 0.75 * L(dq, i + 1);
        L(bl, i) = mn2(mx3(0.0, pmp_1, lac_1), mx2(L(bl, i), mn3(0.0, pmp_1, lac_1)));
      }
    }
  }
  if (iord == 9 || iord == 13)
    fv3o_pert_ppm(ie1 - is1 + 1, &Q(is1), &L(bl, is1), &L(br, is1), 0);

  if (edge) {
    if (is == 1) {
      L(bl, 0) = S14 * L(dm, -1) + S11 * (Q(-1) - Q(0));
      double xt = 0.5 * (((2.0 * DXA(0) + DXA(-1)) * Q(0) - DXA(0) * Q(-1)) / (DXA(-1) + DXA(0)) +
                         ((2.0 * DXA(1) + DXA(2)) * Q(1) - DXA(1) * Q(2)) / (DXA(1) + DXA(2)));
      xt = mx2(xt, mn4(Q(-1), Q(0), Q(1), Q(2)));
      xt = mn2(xt, mx4(Q(-1), Q(0), Q(1), Q(2)));
      L(br, 0) = xt - Q(0);
      L(bl, 1) = xt - Q(1);
      xt = S15 * Q(1) + S11 * Q(2) - S14 * L(dm, 2);
      L(br, 1) = xt - Q(1);
      L(bl, 2) = xt - Q(2);
      L(br, 2) = L(al, 3) - Q(2);
      fv3o_pert_ppm(3, &Q(0), &L(bl, 0), &L(br, 0), 1);
    }
    if (ie + 1 == npx) {
      L(bl, npx - 2) = L(al, npx - 2) - Q(npx - 2);
      double xt = S15 * Q(npx - 1) + S11 * Q(npx - 2) + S14 * L(dm, npx - 2);
      L(br, npx - 2) = xt - Q(npx - 2);
      L(bl, npx - 1) = xt - Q(npx - 1);
      xt = 0.5 * (((2.0 * DXA(npx - 1) + DXA(npx - 2)) * Q(npx - 1) - DXA(npx - 1) * Q(npx - 2)) /
                      (DXA(npx - 2) + DXA(npx - 1)) +
                  ((2.0 * DXA(npx) + DXA(npx + 1)) * Q(npx) - DXA(npx) * Q(npx + 1)) /
                      (DXA(npx) + DXA(npx + 1)));
      xt = mx2(xt, mn4(Q(npx - 2), Q(npx - 1), Q(npx), Q(npx + 1)));
      xt = mn2(xt, mx4(Q(npx - 2), Q(npx - 1), Q(npx), Q(npx + 1)));
      L(br, npx - 1) = xt - Q(npx - 1);
      L(bl, npx) = xt - Q(npx);
      L(br, npx) = S11 * (Q(npx + 1) - Q(npx)) - S14 * L(dm, npx + 1);
      fv3o_pert_ppm(3, &Q(npx - 2), &L(bl, npx - 2), &L(br, npx - 2), 1);
    }
  }

  for (int i = is; i <= ie + 1; i++) {
    double c = c1[i - is];
    if (c > 0.0)
      flux1[i - is] = Q(i - 1) + (1.0 - c) * (L(br, i - 1) - c * (L(bl, i - 1) + L(br, i - 1)));
    else
      flux1[i - is] = Q(i) + (1.0 + c) * (L(bl, i) + c * (L(bl, i) + L(br, i)));
  }
#undef Q
#undef DXA
#undef L
  free(buf);
  return 0;
}

int fv3o_xppm(double *flux, const double *q, const double *c, int iord, int is, int ie, int isd,
              int ied, int jfirst, int jlast, int jsd, int jed, int npx, int npy,
              const double *dxa, int bounded_domain, int grid_type, double lim_fac) {
  (void)npy;
  (void)jed;
  (void)lim_fac; /* lim_fac only enters the non-monotone orders (iord < 7) */
  const int niq = ied - isd + 1, nic = ie - is + 2;
  for (int j = jfirst; j <= jlast; j++) {
    const double *q1 = &E2(q, isd, j, isd, jfirst, niq);
    const double *c1 = &E2(c, is, j, is, jfirst, nic);
    const double *d1 = &E2(dxa, isd, j, isd, jsd, niq);
    double *f1 = &E2(flux, is, j, is, jfirst, nic);
    if (ppm_line(f1, q1, c1, d1, iord, is, ie, isd, ied, npx, bounded_domain, grid_type))
      return 1;
  }
  return 0;
}

int fv3o_yppm(double *flux, const double *q, const double *c, int jord, int ifirst, int ilast,
              int isd, int ied, int js, int je, int jsd, int jed, int npx, int npy,
              const double *dya, int bounded_domain, int grid_type, double lim_fac) {
  (void)npx;
  (void)lim_fac;
  const int niq = ilast - ifirst + 1, nid = ied - isd + 1;
  const int nj = jed - jsd + 1, nf = je - js + 2;
  double *q1 = (double *)malloc(sizeof(double) * nj);
  double *d1 = (double *)malloc(sizeof(double) * nj);
  double *c1 = (double *)malloc(sizeof(double) * nf);
  double *f1 = (double *)malloc(sizeof(double) * nf);
  int rc = 0;
  for (int i = ifirst; i <= ilast && !rc; i++) {
    for (int j = jsd; j <= jed; j++) {
      q1[j - jsd] = E2(q, i, j, ifirst, jsd, niq);
      d1[j - jsd] = E2(dya, i, j, isd, jsd, nid);
    }
    for (int j = js; j <= je + 1; j++) c1[j - js] = E2(c, i, j, isd, js, nid);
    rc = ppm_line(f1, q1, c1, d1, jord, js, je, jsd, jed, npy, bounded_domain, grid_type);
    for (int j = js; j <= je + 1; j++) E2(flux, i, j, ifirst, js, niq) = f1[j - js];
  }
  free(q1);
  free(d1);
  free(c1);
  free(f1);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* copy_corners (tp_core.F90): fill the 3x3 halo block at each cube corner
 * from the adjacent halo strip, rotated for an x (dir=1) or y (dir=2) sweep. */
void fv3o_copy_corners(double *q, int npx, int npy, int dir, const fv3o_bounds *bd,
                       const fv3o_grid *g) {
  if (g->bounded_domain) return;
  const int isd = bd->isd, jsd = bd->jsd, ni = bd->ied - bd->isd + 1;
#define QC(i, j) E2(q, i, j, isd, jsd, ni)
  if (dir == 1) {
    if (g->sw_corner)
      for (int j = 1 - NG; j <= 0; j++)
        for (int i = 1 - NG; i <= 0; i++) QC(i, j) = QC(j, 1 - i);
    if (g->se_corner)
      for (int j = 1 - NG; j <= 0; j++)
        for (int i = npx; i <= npx + NG - 1; i++) QC(i, j) = QC(npy - j, i - npx + 1);
    if (g->ne_corner)
      for (int j = npy; j <= npy + NG - 1; j++)
        for (int i = npx; i <= npx + NG - 1; i++) QC(i, j) = QC(j, 2 * npx - 1 - i);
    if (g->nw_corner)
      for (int j = npy; j <= npy + NG - 1; j++)
        for (int i = 1 - NG; i <= 0; i++) QC(i, j) = QC(npy - j, i - 1 + npx);
  } else if (dir == 2) {
    if (g->sw_corner)
      for (int j = 1 - NG; j <= 0; j++)
        for (int i = 1 - NG; i <= 0; i++) QC(i, j) = QC(1 - j, i);
    if (g->se_corner)
      for (int j = 1 - NG; j <= 0; j++)
        for (int i = npx; i <= npx + NG - 1; i++) QC(i, j) = QC(npy + j - 1, npx - i);
    if (g->ne_corner)
      for (int j = npy; j <= npy + NG - 1; j++)
        for (int i = npx; i <= npx + NG - 1; i++) QC(i, j) = QC(2 * npy - 1 - j, i);
    if (g->nw_corner)
      for (int j = npy; j <= npy + NG - 1; j++)
        for (int i = 1 - NG; i <= 0; i++) QC(i, j) = QC(j + 1 - npx, npy - i);
  }
#undef QC
}

/* ------------------------------------------------------------------------ */
/* deln_flux (tp_core.F90): del-2(nord+1) diffusive flux added to fx, fy.
 * nord 0..2 with a 3-cell halo.  mass == NULL is the unweighted form. */
void fv3o_deln_flux(int nord, int npx, int npy, double damp, const double *q, double *fx,
                    double *fy, const fv3o_grid *g, const fv3o_bounds *bd, const double *mass) {
  const int is = bd->is, ie = bd->ie, js = bd->js, je = bd->je;
  const int isd = bd->isd, ied = bd->ied, jsd = bd->jsd, jed = bd->jed;
  const int nid = ied - isd + 1, njd = jed - jsd + 1;
  const int nfx = ie - is + 2, nfy = ie - is + 1;
  double *d2 = (double *)malloc(sizeof(double) * nid * njd);
  double *fx2 = (double *)malloc(sizeof(double) * (nid + 1) * njd); /* (isd:ied+1, jsd:jed) */
  double *fy2 = (double *)malloc(sizeof(double) * nid * (njd + 1)); /* (isd:ied, jsd:jed+1) */
  for (int t = 0; t < nid * njd; t++) d2[t] = NAN;
  for (int t = 0; t < (nid + 1) * njd; t++) fx2[t] = NAN;
  for (int t = 0; t < nid * (njd + 1); t++) fy2[t] = NAN;
#define D2(i, j) E2(d2, i, j, isd, jsd, nid)
#define FX2(i, j) E2(fx2, i, j, isd, jsd, nid + 1)
#define FY2(i, j) E2(fy2, i, j, isd, jsd, nid)
#define DEL6U(i, j) E2(g->del6_u, i, j, isd, jsd, nid)
#define DEL6V(i, j) E2(g->del6_v, i, j, isd, jsd, nid + 1)
  const int i1 = is - 1 - nord, i2 = ie + 1 + nord, j1 = js - 1 - nord, j2 = je + 1 + nord;
  for (int j = j1; j <= j2; j++)
    for (int i = i1; i <= i2; i++)
      D2(i, j) = mass == NULL ? damp * E2(q, i, j, isd, jsd, nid) : E2(q, i, j, isd, jsd, nid);

  if (nord > 0) fv3o_copy_corners(d2, npx, npy, 1, bd, g);
  for (int j = js - nord; j <= je + nord; j++)
    for (int i = is - nord; i <= ie + nord + 1; i++) FX2(i, j) = DEL6V(i, j) * (D2(i - 1, j) - D2(i, j));

  if (nord > 0) fv3o_copy_corners(d2, npx, npy, 2, bd, g);
  for (int j = js - nord; j <= je + nord + 1; j++)
    for (int i = is - nord; i <= ie + nord; i++) FY2(i, j) = DEL6U(i, j) * (D2(i, j - 1) - D2(i, j));

  for (int n = 1; n <= nord; n++) {
    const int nt = nord - n;
    for (int j = js - nt - 1; j <= je + nt + 1; j++)
      for (int i = is - nt - 1; i <= ie + nt + 1; i++)
        D2(i, j) = (FX2(i, j) - FX2(i + 1, j) + FY2(i, j) - FY2(i, j + 1)) *
                   E2(g->rarea, i, j, isd, jsd, nid);
    fv3o_copy_corners(d2, npx, npy, 1, bd, g);
    for (int j = js - nt; j <= je + nt; j++)
      for (int i = is - nt; i <= ie + nt + 1; i++) FX2(i, j) = DEL6V(i, j) * (D2(i, j) - D2(i - 1, j));
    fv3o_copy_corners(d2, npx, npy, 2, bd, g);
    for (int j = js - nt; j <= je + nt + 1; j++)
      for (int i = is - nt; i <= ie + nt; i++) FY2(i, j) = DEL6U(i, j) * (D2(i, j) - D2(i, j - 1));
  }

  if (mass != NULL) {
    const double damp2 = 0.5 * damp;
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie + 1; i++)
        E2(fx, i, j, is, js, nfx) = E2(fx, i, j, is, js, nfx) +
            damp2 * (E2(mass, i - 1, j, isd, jsd, nid) + E2(mass, i, j, isd, jsd, nid)) * FX2(i, j);
    for (int j = js; j <= je + 1; j++)
      for (int i = is; i <= ie; i++)
        E2(fy, i, j, is, js, nfy) = E2(fy, i, j, is, js, nfy) +
            damp2 * (E2(mass, i, j - 1, isd, jsd, nid) + E2(mass, i, j, isd, jsd, nid)) * FY2(i, j);
  } else {
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie + 1; i++) E2(fx, i, j, is, js, nfx) = E2(fx, i, j, is, js, nfx) + FX2(i, j);
    for (int j = js; j <= je + 1; j++)
      for (int i = is; i <= ie; i++) E2(fy, i, j, is, js, nfy) = E2(fy, i, j, is, js, nfy) + FY2(i, j);
  }
#undef D2
#undef FX2
#undef FY2
#undef DEL6U
#undef DEL6V
  free(d2);
  free(fx2);
  free(fy2);
}

/* ------------------------------------------------------------------------ */
/* fv_tp_2d (tp_core.F90): Lin-Rood transport fluxes of one 2-D scalar. */
int fv3o_fv_tp_2d(double *q, const double *crx, const double *cry, int npx, int npy, int hord,
                  double *fx, double *fy, const double *xfx, const double *yfx,
                  const fv3o_grid *g, const fv3o_bounds *bd, const double *ra_x,
                  const double *ra_y, double lim_fac, const double *mfx, const double *mfy,
                  const double *mass, const int *nord, const double *damp_c) {
  const int is = bd->is, ie = bd->ie, js = bd->js, je = bd->je;
  const int isd = bd->isd, ied = bd->ied, jsd = bd->jsd, jed = bd->jed;
  const int nid = ied - isd + 1, njd = jed - jsd + 1;
  const int nix = ie - is + 2, niy = ie - is + 1;
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  int rc = 0;

  double *q_i = (double *)malloc(sizeof(double) * nid * (je - js + 1));    /* (isd:ied, js:je) */
  double *q_j = (double *)malloc(sizeof(double) * niy * njd);              /* (is:ie, jsd:jed) */
  double *fx2 = (double *)malloc(sizeof(double) * nix * njd);              /* (is:ie+1, jsd:jed) */
  double *fy2 = (double *)malloc(sizeof(double) * nid * (je - js + 2));    /* (isd:ied, js:je+1) */
  double *fyy = (double *)malloc(sizeof(double) * nid * (je - js + 2));

  if (!g->bounded_domain) fv3o_copy_corners(q, npx, npy, 2, bd, g);
  rc |= fv3o_yppm(fy2, q, cry, ord_in, isd, ied, isd, ied, js, je, jsd, jed, npx, npy, g->dya,
                  g->bounded_domain, g->grid_type, lim_fac);
  for (int j = js; j <= je + 1; j++)
    for (int i = isd; i <= ied; i++)
      E2(fyy, i, j, isd, js, nid) = E2(yfx, i, j, isd, js, nid) * E2(fy2, i, j, isd, js, nid);
  for (int j = js; j <= je; j++)
    for (int i = isd; i <= ied; i++)
      E2(q_i, i, j, isd, js, nid) =
          (E2(q, i, j, isd, jsd, nid) * E2(g->area, i, j, isd, jsd, nid) +
           E2(fyy, i, j, isd, js, nid) - E2(fyy, i, j + 1, isd, js, nid)) /
          E2(ra_y, i, j, isd, js, nid);

  rc |= fv3o_xppm(fx, q_i, &E2(crx, is, js, is, jsd, nix), ord_ou, is, ie, isd, ied, js, je,
                  jsd, jed, npx, npy, g->dxa, g->bounded_domain, g->grid_type, lim_fac);

  if (!g->bounded_domain) fv3o_copy_corners(q, npx, npy, 1, bd, g);
  rc |= fv3o_xppm(fx2, q, crx, ord_in, is, ie, isd, ied, jsd, jed, jsd, jed, npx, npy, g->dxa,
                  g->bounded_domain, g->grid_type, lim_fac);
  for (int j = jsd; j <= jed; j++) {
    for (int i = is; i <= ie; i++) {
      double f1 = E2(xfx, i, j, is, jsd, nix) * E2(fx2, i, j, is, jsd, nix);
      double f2 = E2(xfx, i + 1, j, is, jsd, nix) * E2(fx2, i + 1, j, is, jsd, nix);
      E2(q_j, i, j, is, jsd, niy) =
          (E2(q, i, j, isd, jsd, nid) * E2(g->area, i, j, isd, jsd, nid) + f1 - f2) /
          E2(ra_x, i, j, is, jsd, niy);
    }
  }
  rc |= fv3o_yppm(fy, q_j, cry, ord_ou, is, ie, isd, ied, js, je, jsd, jed, npx, npy, g->dya,
                  g->bounded_domain, g->grid_type, lim_fac);

  if (mfx != NULL && mfy != NULL) {
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie + 1; i++)
        E2(fx, i, j, is, js, nix) = 0.5 * (E2(fx, i, j, is, js, nix) + E2(fx2, i, j, is, jsd, nix)) *
                                    E2(mfx, i, j, is, js, nix);
    for (int j = js; j <= je + 1; j++)
      for (int i = is; i <= ie; i++)
        E2(fy, i, j, is, js, niy) = 0.5 * (E2(fy, i, j, is, js, niy) + E2(fy2, i, j, isd, js, nid)) *
                                    E2(mfy, i, j, is, js, niy);
    if (nord != NULL && damp_c != NULL && mass != NULL && *damp_c > 1.0e-4) {
      double damp = pow(*damp_c * g->da_min, (double)(*nord + 1));
      fv3o_deln_flux(*nord, npx, npy, damp, q, fx, fy, g, bd, mass);
    }
  } else {
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie + 1; i++)
        E2(fx, i, j, is, js, nix) = 0.5 * (E2(fx, i, j, is, js, nix) + E2(fx2, i, j, is, jsd, nix)) *
                                    E2(xfx, i, j, is, jsd, nix);
    for (int j = js; j <= je + 1; j++)
      for (int i = is; i <= ie; i++)
        E2(fy, i, j, is, js, niy) = 0.5 * (E2(fy, i, j, is, js, niy) + E2(fy2, i, j, isd, js, nid)) *
                                    E2(yfx, i, j, isd, js, nid);
    if (nord != NULL && damp_c != NULL && *damp_c > 1.0e-4) {
      double damp = pow(*damp_c * g->da_min, (double)(*nord + 1));
      fv3o_deln_flux(*nord, npx, npy, damp, q, fx, fy, g, bd, NULL);
    }
  }
  free(q_i);
  free(q_j);
  free(fx2);
  free(fy2);
  free(fyy);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Shared pieces of tracer_2d / tracer_2d_1L (fv_tracer2d.F90). */

typedef struct {
  int nid, njd, nix, niy;       /* leading extents */
  size_t s2, sx, sy, smx, smy;  /* per-level sizes of q/dp1, cx, cy, mfx, mfy */
} tr_dims;

static tr_dims tr_dims_of(const fv3o_bounds *bd) {
  tr_dims d;
  d.nid = bd->ied - bd->isd + 1;
  d.njd = bd->jed - bd->jsd + 1;
  d.nix = bd->ie - bd->is + 2;
  d.niy = bd->ie - bd->is + 1;
  d.s2 = (size_t)d.nid * d.njd;
  d.sx = (size_t)d.nix * d.njd;
  d.sy = (size_t)d.nid * (bd->je - bd->js + 2);
  d.smx = (size_t)d.nix * (bd->je - bd->js + 1);
  d.smy = (size_t)d.niy * (bd->je - bd->js + 2);
  return d;
}

/* Upwind area fluxes xfx, yfx and the per-level Courant maximum. */
static void tr_area_fluxes(const double *cx, const double *cy, double *xfx, double *yfx,
                           double *cmax, const fv3o_grid *g, const fv3o_bounds *bd, int npz) {
  const int is = bd->is, ie = bd->ie, js = bd->js, je = bd->je;
  const int isd = bd->isd, ied = bd->ied, jsd = bd->jsd, jed = bd->jed;
  const tr_dims d = tr_dims_of(bd);
  const size_t s2 = d.s2;
  for (int k = 1; k <= npz; k++) {
    const double *cxk = cx + (k - 1) * d.sx, *cyk = cy + (k - 1) * d.sy;
    double *xk = xfx + (k - 1) * d.sx, *yk = yfx + (k - 1) * d.sy;
    for (int j = jsd; j <= jed; j++)
      for (int i = is; i <= ie + 1; i++) {
        double c = E2(cxk, i, j, is, jsd, d.nix);
        if (c > 0.0)
          E2(xk, i, j, is, jsd, d.nix) = c * E2(g->dxa, i - 1, j, isd, jsd, d.nid) *
                                          E2(g->dy, i, j, isd, jsd, d.nid + 1) *
                                          E2(g->sin_sg + 2 * s2, i - 1, j, isd, jsd, d.nid);
        else
          E2(xk, i, j, is, jsd, d.nix) = c * E2(g->dxa, i, j, isd, jsd, d.nid) *
                                          E2(g->dy, i, j, isd, jsd, d.nid + 1) *
                                          E2(g->sin_sg + 0 * s2, i, j, isd, jsd, d.nid);
      }
    for (int j = js; j <= je + 1; j++)
      for (int i = isd; i <= ied; i++) {
        double c = E2(cyk, i, j, isd, js, d.nid);
        if (c > 0.0)
          E2(yk, i, j, isd, js, d.nid) = c * E2(g->dya, i, j - 1, isd, jsd, d.nid) *
                                          E2(g->dx, i, j, isd, jsd, d.nid) *
                                          E2(g->sin_sg + 3 * s2, i, j - 1, isd, jsd, d.nid);
        else
          E2(yk, i, j, isd, js, d.nid) = c * E2(g->dya, i, j, isd, jsd, d.nid) *
                                          E2(g->dx, i, j, isd, jsd, d.nid) *
                                          E2(g->sin_sg + 1 * s2, i, j, isd, jsd, d.nid);
      }
    double cm = 0.0;
    if (k < npz / 6) {
      for (int j = js; j <= je; j++)
        for (int i = is; i <= ie; i++)
          cm = mx3(cm, fabs(E2(cxk, i, j, is, jsd, d.nix)), fabs(E2(cyk, i, j, isd, js, d.nid)));
    } else {
      for (int j = js; j <= je; j++)
        for (int i = is; i <= ie; i++)
          cm = mx2(cm, mx2(fabs(E2(cxk, i, j, is, jsd, d.nix)), fabs(E2(cyk, i, j, isd, js, d.nid))) +
                           1.0 - E2(g->sin_sg + 4 * s2, i, j, isd, jsd, d.nid));
    }
    cmax[k - 1] = cm;
  }
}

/* In-place scaling of cx, xfx, mfx, cy, yfx, mfy of level k by frac. */
static void tr_scale_level(double *cx, double *xfx, double *mfx, double *cy, double *yfx,
                           double *mfy, const tr_dims *d, int k, double frac) {
  for (size_t t = 0; t < d->sx; t++) {
    cx[(k - 1) * d->sx + t] = cx[(k - 1) * d->sx + t] * frac;
    xfx[(k - 1) * d->sx + t] = xfx[(k - 1) * d->sx + t] * frac;
  }
  for (size_t t = 0; t < d->smx; t++) mfx[(k - 1) * d->smx + t] = mfx[(k - 1) * d->smx + t] * frac;
  for (size_t t = 0; t < d->sy; t++) {
    cy[(k - 1) * d->sy + t] = cy[(k - 1) * d->sy + t] * frac;
    yfx[(k - 1) * d->sy + t] = yfx[(k - 1) * d->sy + t] * frac;
  }
  for (size_t t = 0; t < d->smy; t++) mfy[(k - 1) * d->smy + t] = mfy[(k - 1) * d->smy + t] * frac;
}

/* ra_x (is:ie, jsd:jed) and ra_y (isd:ied, js:je) of one level. */
static void tr_ra(const double *xk, const double *yk, double *ra_x, double *ra_y,
                  const fv3o_grid *g, const fv3o_bounds *bd) {
  const tr_dims d = tr_dims_of(bd);
  for (int j = bd->jsd; j <= bd->jed; j++)
    for (int i = bd->is; i <= bd->ie; i++)
      E2(ra_x, i, j, bd->is, bd->jsd, d.niy) = E2(g->area, i, j, bd->isd, bd->jsd, d.nid) +
                                               E2(xk, i, j, bd->is, bd->jsd, d.nix) -
                                               E2(xk, i + 1, j, bd->is, bd->jsd, d.nix);
  for (int j = bd->js; j <= bd->je; j++)
    for (int i = bd->isd; i <= bd->ied; i++)
      E2(ra_y, i, j, bd->isd, bd->js, d.nid) = E2(g->area, i, j, bd->isd, bd->jsd, d.nid) +
                                               E2(yk, i, j, bd->isd, bd->js, d.nid) -
                                               E2(yk, i, j + 1, bd->isd, bd->js, d.nid);
}

/* dp2 (is:ie, js:je) of one level. */
static void tr_dp2(const double *dp1k, const double *mfxk, const double *mfyk, double *dp2,
                   const fv3o_grid *g, const fv3o_bounds *bd) {
  const tr_dims d = tr_dims_of(bd);
  for (int j = bd->js; j <= bd->je; j++)
    for (int i = bd->is; i <= bd->ie; i++)
      E2(dp2, i, j, bd->is, bd->js, d.niy) =
          E2(dp1k, i, j, bd->isd, bd->jsd, d.nid) +
          ((E2(mfxk, i, j, bd->is, bd->js, d.nix) - E2(mfxk, i + 1, j, bd->is, bd->js, d.nix)) +
           (E2(mfyk, i, j, bd->is, bd->js, d.niy) - E2(mfyk, i, j + 1, bd->is, bd->js, d.niy))) *
              E2(g->rarea, i, j, bd->isd, bd->jsd, d.nid);
}

/* Flux-form update q_out = (q_in*dp1 + div(F)*rarea)/dp2 over the interior. */
static void tr_update(double *q_out, const double *q_in, const double *dp1k, const double *fx,
                      const double *fy, const double *dp2, const fv3o_grid *g,
                      const fv3o_bounds *bd) {
  const tr_dims d = tr_dims_of(bd);
  for (int j = bd->js; j <= bd->je; j++)
    for (int i = bd->is; i <= bd->ie; i++)
      E2(q_out, i, j, bd->isd, bd->jsd, d.nid) =
          (E2(q_in, i, j, bd->isd, bd->jsd, d.nid) * E2(dp1k, i, j, bd->isd, bd->jsd, d.nid) +
           (E2(fx, i, j, bd->is, bd->js, d.nix) - E2(fx, i + 1, j, bd->is, bd->js, d.nix) +
            E2(fy, i, j, bd->is, bd->js, d.niy) - E2(fy, i, j + 1, bd->is, bd->js, d.niy)) *
               E2(g->rarea, i, j, bd->isd, bd->jsd, d.nid)) /
          E2(dp2, i, j, bd->is, bd->js, d.niy);
}

typedef struct {
  tr_dims d;
  size_t tq, tdp, tcx, tcy, tmx, tmy; /* per-tile strides */
  int npz;
} tr_tiles;

static tr_tiles tr_tiles_of(const fv3o_bounds *bd, int npz, int nq) {
  tr_tiles t;
  t.d = tr_dims_of(bd);
  t.tq = t.d.s2 * npz * nq;
  t.tdp = t.d.s2 * npz;
  t.tcx = t.d.sx * npz;
  t.tcy = t.d.sy * npz;
  t.tmx = t.d.smx * npz;
  t.tmy = t.d.smy * npz;
  t.npz = npz;
  return t;
}

/* One level of tracer_2d_1L: nsplt sub-cycles of every tile and tracer.
 * Work buffers: ra_x, ra_y, dp2 (per tile), qn2 (per tile and tracer), fx, fy. */
typedef struct {
  double *qn2, *dp2, *ra_x, *ra_y, *fx, *fy;
} tr_work;

static int tr_alloc(tr_work *w, const fv3o_bounds *bd, int ntiles, int nq) {
  const tr_dims d = tr_dims_of(bd);
  const size_t nj = (size_t)(bd->je - bd->js + 1);
  w->qn2 = (double *)malloc(sizeof(double) * d.s2 * nq * ntiles);
  w->dp2 = (double *)malloc(sizeof(double) * d.niy * nj * ntiles);
  w->ra_x = (double *)malloc(sizeof(double) * d.niy * d.njd * ntiles);
  w->ra_y = (double *)malloc(sizeof(double) * d.nid * nj * ntiles);
  w->fx = (double *)malloc(sizeof(double) * d.smx);
  w->fy = (double *)malloc(sizeof(double) * d.smy);
  return w->qn2 && w->dp2 && w->ra_x && w->ra_y && w->fx && w->fy ? 0 : 1;
}

static void tr_free(tr_work *w) {
  free(w->qn2);
  free(w->dp2);
  free(w->ra_x);
  free(w->ra_y);
  free(w->fx);
  free(w->fy);
}

static int tr_level_1L(int k, int nsplt, double *q, double *dp1, double *mfx, double *mfy,
                       double *cx, double *cy, const double *xfx, const double *yfx,
                       const fv3o_grid *g, const fv3o_bounds *bd, const tr_tiles *T, int ntiles,
                       int npx, int npy, int nq, int hord, double lim_fac, fv3o_halo_fn halo,
                       void *halo_ctx, tr_work *w) {
  const tr_dims d = T->d;
  const size_t sdp2 = (size_t)d.niy * (bd->je - bd->js + 1);
  const size_t srax = (size_t)d.niy * d.njd, sray = (size_t)d.nid * (bd->je - bd->js + 1);
  int rc = 0;
  for (int t = 0; t < ntiles; t++)
    tr_ra(xfx + t * T->tcx + (k - 1) * d.sx, yfx + t * T->tcy + (k - 1) * d.sy, w->ra_x + t * srax,
          w->ra_y + t * sray, &g[t], bd);
  for (int it = 1; it <= nsplt && !rc; it++) {
    for (int t = 0; t < ntiles; t++)
      tr_dp2(dp1 + t * T->tdp + (k - 1) * d.s2, mfx + t * T->tmx + (k - 1) * d.smx,
             mfy + t * T->tmy + (k - 1) * d.smy, w->dp2 + t * sdp2, &g[t], bd);
    for (int t = 0; t < ntiles && !rc; t++) {
      double *dp1k = dp1 + t * T->tdp + (k - 1) * d.s2;
      double *cxk = cx + t * T->tcx + (k - 1) * d.sx, *cyk = cy + t * T->tcy + (k - 1) * d.sy;
      const double *xk = xfx + t * T->tcx + (k - 1) * d.sx, *yk = yfx + t * T->tcy + (k - 1) * d.sy;
      double *mxk = mfx + t * T->tmx + (k - 1) * d.smx, *myk = mfy + t * T->tmy + (k - 1) * d.smy;
      for (int iq = 1; iq <= nq && !rc; iq++) {
        double *qk = q + t * T->tq + ((size_t)(iq - 1) * T->npz + (k - 1)) * d.s2;
        double *qn = w->qn2 + ((size_t)t * nq + (iq - 1)) * d.s2;
        double *work = nsplt != 1 ? qn : qk;
        if (nsplt != 1 && it == 1) memcpy(qn, qk, sizeof(double) * d.s2);
        rc |= fv3o_fv_tp_2d(work, cxk, cyk, npx, npy, hord, w->fx, w->fy, xk, yk, &g[t], bd,
                            w->ra_x + t * srax, w->ra_y + t * sray, lim_fac, mxk, myk, NULL, NULL,
                            NULL);
        tr_update(it < nsplt ? qn : qk, work, dp1k, w->fx, w->fy, w->dp2 + t * sdp2, &g[t], bd);
      }
    }
    if (it < nsplt) {
      for (int t = 0; t < ntiles; t++) {
        double *dp1k = dp1 + t * T->tdp + (k - 1) * d.s2;
        for (int j = bd->js; j <= bd->je; j++)
          for (int i = bd->is; i <= bd->ie; i++)
            E2(dp1k, i, j, bd->isd, bd->jsd, d.nid) = E2(w->dp2 + t * sdp2, i, j, bd->is, bd->js, d.niy);
      }
      if (halo) halo(halo_ctx, w->qn2, nq, k); /* mpp_update_domains(qn2) */
    }
  }
  return rc;
}

int fv3o_tracer_2d_1L(double *q, double *dp1, double *mfx, double *mfy, double *cx,
                      double *cy, const fv3o_grid *g, const fv3o_bounds *bd, int ntiles,
                      int npx, int npy, int npz, int nq, int hord, int q_split, double dt,
                      int id_divg, int nord_tr, double trdm, double lim_fac,
                      fv3o_halo_fn halo, void *halo_ctx, int *nsplit_out) {
  (void)q_split; /* per-level sub-cycling always follows cmax(k) */
  (void)dt;
  (void)id_divg;
  (void)nord_tr;
  (void)trdm;
  const tr_tiles T = tr_tiles_of(bd, npz, nq);
  const tr_dims d = T.d;
  int rc = 0;
  double *xfx = (double *)malloc(sizeof(double) * T.tcx * ntiles);
  double *yfx = (double *)malloc(sizeof(double) * T.tcy * ntiles);
  double *cmax = (double *)calloc((size_t)npz * ntiles, sizeof(double));
  double *cm = (double *)calloc(npz, sizeof(double));

  for (int t = 0; t < ntiles; t++)
    tr_area_fluxes(cx + t * T.tcx, cy + t * T.tcy, xfx + t * T.tcx, yfx + t * T.tcy,
                   cmax + (size_t)t * npz, &g[t], bd, npz);
  for (int k = 0; k < npz; k++) { /* mp_reduce_max over tiles */
    cm[k] = cmax[k];
    for (int t = 1; t < ntiles; t++) cm[k] = mx2(cm[k], cmax[(size_t)t * npz + k]);
  }
  for (int k = 1; k <= npz; k++) {
    int nsplt = (int)(1.0 + cm[k - 1]);
    if (nsplit_out) nsplit_out[k - 1] = nsplt;
    if (nsplt > 1) {
      double frac = 1.0 / (double)nsplt;
      for (int t = 0; t < ntiles; t++)
        tr_scale_level(cx + t * T.tcx, xfx + t * T.tcx, mfx + t * T.tmx, cy + t * T.tcy,
                       yfx + t * T.tcy, mfy + t * T.tmy, &d, k, frac);
    }
  }
  if (halo) halo(halo_ctx, q, npz * nq, 0); /* complete_group_halo_update(q_pack) */

  /* Levels are independent (upstream's OpenMP loop over k).  A halo callback
   * (six tiles) runs on the calling thread, so that case stays serial. */
#pragma omp parallel if (halo == NULL) reduction(| : rc)
  {
    tr_work w;
    if (tr_alloc(&w, bd, ntiles, nq)) {
      rc |= 1;
    } else {
#pragma omp for schedule(dynamic, 1)
      for (int k = 1; k <= npz; k++)
        rc |= tr_level_1L(k, (int)(1.0 + cm[k - 1]), q, dp1, mfx, mfy, cx, cy, xfx, yfx, g, bd, &T,
                          ntiles, npx, npy, nq, hord, lim_fac, halo, halo_ctx, &w);
    }
    tr_free(&w);
  }
  free(xfx);
  free(yfx);
  free(cmax);
  free(cm);
  return rc;
}

int fv3o_tracer_2d(double *q, double *dp1, double *mfx, double *mfy, double *cx, double *cy,
                   const fv3o_grid *g, const fv3o_bounds *bd, int ntiles, int npx, int npy,
                   int npz, int nq, int hord, int q_split, double dt, int id_divg,
                   int nord_tr, double trdm, double lim_fac, fv3o_halo_fn halo,
                   void *halo_ctx, int *nsplit_out) {
  const tr_tiles T = tr_tiles_of(bd, npz, nq);
  const tr_dims d = T.d;
  int rc = 0;
  double *xfx = (double *)malloc(sizeof(double) * T.tcx * ntiles);
  double *yfx = (double *)malloc(sizeof(double) * T.tcy * ntiles);
  double *cmax = (double *)calloc((size_t)npz * ntiles, sizeof(double));
  double *dp2 = (double *)malloc(sizeof(double) * d.niy * (bd->je - bd->js + 1));
  double *ra_x = (double *)malloc(sizeof(double) * d.niy * d.njd);
  double *ra_y = (double *)malloc(sizeof(double) * d.nid * (bd->je - bd->js + 1));
  double *fx = (double *)malloc(sizeof(double) * d.smx);
  double *fy = (double *)malloc(sizeof(double) * d.smy);

  for (int t = 0; t < ntiles; t++)
    tr_area_fluxes(cx + t * T.tcx, cy + t * T.tcy, xfx + t * T.tcx, yfx + t * T.tcy,
                   cmax + (size_t)t * npz, &g[t], bd, npz);
  int nsplt;
  if (q_split == 0) {
    for (int k = 0; k < npz; k++) /* mp_reduce_max over tiles */
      for (int t = 1; t < ntiles; t++) cmax[k] = mx2(cmax[k], cmax[(size_t)t * npz + k]);
    double c_global = cmax[0];
    if (npz != 1)
      for (int k = 2; k <= npz; k++) c_global = mx2(cmax[k - 1], c_global);
    nsplt = (int)(1.0 + c_global);
  } else {
    nsplt = q_split;
  }
  if (nsplit_out)
    for (int k = 0; k < npz; k++) nsplit_out[k] = nsplt;
  const double frac = 1.0 / (double)nsplt;
  if (nsplt != 1)
    for (int t = 0; t < ntiles; t++)
      for (int k = 1; k <= npz; k++)
        tr_scale_level(cx + t * T.tcx, xfx + t * T.tcx, mfx + t * T.tmx, cy + t * T.tcy,
                       yfx + t * T.tcy, mfy + t * T.tmy, &d, k, frac);

  for (int it = 1; it <= nsplt && !rc; it++) {
    if (halo) halo(halo_ctx, q, npz * nq, 0); /* complete_group_halo_update(q_pack) */
    for (int t = 0; t < ntiles && !rc; t++) {
      for (int k = 1; k <= npz && !rc; k++) {
        double *dp1k = dp1 + t * T.tdp + (k - 1) * d.s2;
        double *cxk = cx + t * T.tcx + (k - 1) * d.sx, *cyk = cy + t * T.tcy + (k - 1) * d.sy;
        double *xk = xfx + t * T.tcx + (k - 1) * d.sx, *yk = yfx + t * T.tcy + (k - 1) * d.sy;
        double *mxk = mfx + t * T.tmx + (k - 1) * d.smx, *myk = mfy + t * T.tmy + (k - 1) * d.smy;
        tr_dp2(dp1k, mxk, myk, dp2, &g[t], bd);
        tr_ra(xk, yk, ra_x, ra_y, &g[t], bd);
        for (int iq = 1; iq <= nq && !rc; iq++) {
          double *qk = q + t * T.tq + ((size_t)(iq - 1) * npz + (k - 1)) * d.s2;
          if (it == 1 && trdm > 1.0e-4)
            rc |= fv3o_fv_tp_2d(qk, cxk, cyk, npx, npy, hord, fx, fy, xk, yk, &g[t], bd, ra_x,
                                ra_y, lim_fac, mxk, myk, dp1k, &nord_tr, &trdm);
          else
            rc |= fv3o_fv_tp_2d(qk, cxk, cyk, npx, npy, hord, fx, fy, xk, yk, &g[t], bd, ra_x,
                                ra_y, lim_fac, mxk, myk, NULL, NULL, NULL);
          tr_update(qk, qk, dp1k, fx, fy, dp2, &g[t], bd);
        }
        if (it != nsplt)
          for (int j = bd->js; j <= bd->je; j++)
            for (int i = bd->is; i <= bd->ie; i++)
              E2(dp1k, i, j, bd->isd, bd->jsd, d.nid) = E2(dp2, i, j, bd->is, bd->js, d.niy);
      }
    }
  }
  if (id_divg > 0 && !rc) {
    const double rdt = 1.0 / (frac * dt);
    for (int t = 0; t < ntiles; t++)
      for (int k = 1; k <= npz; k++) {
        double *dp1k = dp1 + t * T.tdp + (k - 1) * d.s2;
        const double *xk = xfx + t * T.tcx + (k - 1) * d.sx, *yk = yfx + t * T.tcy + (k - 1) * d.sy;
        for (int j = bd->js; j <= bd->je; j++)
          for (int i = bd->is; i <= bd->ie; i++)
            E2(dp1k, i, j, bd->isd, bd->jsd, d.nid) =
                (E2(xk, i + 1, j, bd->is, bd->jsd, d.nix) - E2(xk, i, j, bd->is, bd->jsd, d.nix) +
                 E2(yk, i, j + 1, bd->isd, bd->js, d.nid) - E2(yk, i, j, bd->isd, bd->js, d.nid)) *
                E2(g[t].rarea, i, j, bd->isd, bd->jsd, d.nid) * rdt;
      }
  }
  free(xfx);
  free(yfx);
  free(cmax);
  free(dp2);
  free(ra_x);
  free(ra_y);
  free(fx);
  free(fy);
  return rc;
}

/* OpenMP threads of the oracle's level and row loops (the CPU baseline times
 * one thread and all of them). */
#ifdef _OPENMP
#include <omp.h>
#endif
int fv3o_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}
