/*
 * fv3_oracle_l2e.c -- FP64 CPU restatement of the rest of FV3's
 * Lagrangian_to_Eulerian (upstream fv_mapz.F90; empty submodule,
 * .gitmodules:1-4): the remap of the vertical velocity w (bottom boundary
 * value ws, kord_wz), of the layer thickness delz (as -delz/delp, kord_wz),
 * of the D-grid winds u and v on pressure interfaces averaged to the cell
 * faces (kord_mt), the recomputation of pe, ps, peln, pk and pkz on the
 * Eulerian levels, and a total-energy fix-up that adds the energy the remap
 * lost back as one temperature increment over the whole domain.  The
 * tracers, pt and delp go through fv3o_lagrangian_to_eulerian_tracers.
 * TEST INFRASTRUCTURE ONLY -- the checker for the HIP product.
 *
 * The pressure-variable set (ps, pe, peln, pk, pkz, ptop) is the one of
 * moving_nest/fv_moving_nest.F90:2205-2213; the hybrid-coordinate pressures
 * follow ccpp/data/CCPP_typedefs.F90:1691-1711.  The energy fix-up's exact
 * upstream form is unpinned (absent source): this restatement keeps the
 * total energy sum_cells sum_k (cp T + (ua^2 + va^2)/2) dp area (hydrostatic,
 * T = pt pkz, ua/va the face-averaged D-grid winds) equal before and after
 * the remap when consv = 1.
 */
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "fv3_oracle.h"

/* Neumaier (compensated) summation, in the given order. */
typedef struct {
  double s, c;
} ksum;
static void kadd(ksum *k, double x) {
  const double t = k->s + x;
  if (fabs(k->s) >= fabs(x))
    k->c += (k->s - t) + x;
  else
    k->c += (x - t) + k->s;
  k->s = t;
}
static double kval(const ksum *k) { return k->s + k->c; }

/* Column total energy of cell (i, j): sum_k (cp pt pkz + (ua^2 + va^2)/2) dp. */
static double column_te(const fv3o_l2e *a, int t, int i, int j, const double *dp, size_t dps,
                        const double *pkz) {
  const fv3o_bounds *b = &a->bd;
  const size_t nid = (size_t)(b->ied - b->isd + 1), njd = (size_t)(b->jed - b->jsd + 1);
  const size_t s2 = nid * njd, su = nid * (njd + 1), sv = (nid + 1) * njd;
  const size_t ni = (size_t)(b->ie - b->is + 1), nj = (size_t)(b->je - b->js + 1);
  const int km = a->km;
  double te = 0.0;
  for (int k = 1; k <= km; k++) {
    const double *uk = a->u + ((size_t)t * km + (k - 1)) * su;
    const double *vk = a->v + ((size_t)t * km + (k - 1)) * sv;
    const double ua = 0.5 * (uk[(size_t)(j - b->jsd) * nid + (i - b->isd)] +
                             uk[(size_t)(j + 1 - b->jsd) * nid + (i - b->isd)]);
    const double va = 0.5 * (vk[(size_t)(j - b->jsd) * (nid + 1) + (i - b->isd)] +
                             vk[(size_t)(j - b->jsd) * (nid + 1) + (i + 1 - b->isd)]);
    const double ptk = a->pt[((size_t)t * km + (k - 1)) * s2 + (size_t)(j - b->jsd) * nid + (i - b->isd)];
    const double pz = pkz[((size_t)t * km + (k - 1)) * ni * nj + (size_t)(j - b->js) * ni + (i - b->is)];
    const double d = dp[(size_t)(k - 1) * dps];
    te = te + (a->cp * ptk * pz + 0.5 * (ua * ua + va * va)) * d;
  }
  return te;
}

int fv3o_lagrangian_to_eulerian_full(const fv3o_l2e *a) {
  const fv3o_bounds *b = &a->bd;
  const int km = a->km, nt = a->ntiles;
  const int is = b->is, ie = b->ie, js = b->js, je = b->je;
  const int isd = b->isd, jsd = b->jsd;
  const size_t nid = (size_t)(b->ied - isd + 1), njd = (size_t)(b->jed - jsd + 1);
  const size_t s2 = nid * njd, su = nid * (njd + 1), sv = (nid + 1) * njd;
  const size_t ni = (size_t)(ie - is + 1), nj = (size_t)(je - js + 1);
  /* pe: (is-1:ie+1, km+1, js-1:je+1); peln: (is:ie, km+1, js:je);
   * pk: (is:ie, js:je, km+1); pkz, delz: (is:ie, js:je, km) */
  const size_t npi = ni + 2, npj = nj + 2;
  const size_t spe = npi * (km + 1) * npj, spl = ni * (km + 1) * nj;
  const size_t spk = ni * nj * (km + 1), spz = ni * nj * km;
  int rc = 0;
#define PE(t, i, k, j) a->pe[(size_t)(t) * spe + ((size_t)((j) - (js - 1)) * (km + 1) + ((k) - 1)) * npi + ((i) - (is - 1))]
#define DELP(t, i, j, k) a->delp[((size_t)(t) * km + ((k) - 1)) * s2 + (size_t)((j) - jsd) * nid + ((i) - isd)]
#define PS(t, i, j) a->ps[(size_t)(t) * s2 + (size_t)((j) - jsd) * nid + ((i) - isd)]
  /* Lagrangian interfaces on the compute domain and its one-cell ring */
  for (int t = 0; t < nt; t++)
    for (int j = js - 1; j <= je + 1; j++)
      for (int i = is - 1; i <= ie + 1; i++) {
        PE(t, i, 1, j) = a->ptop;
        for (int k = 1; k <= km; k++) PE(t, i, k + 1, j) = PE(t, i, k, j) + DELP(t, i, j, k);
        PS(t, i, j) = PE(t, i, km + 1, j);
      }
  /* column energy before the remap (Lagrangian thickness, pkz as given) */
  double *dpc = (double *)malloc(sizeof(double) * km);
  if (a->te0_2d)
    for (int t = 0; t < nt; t++)
      for (int j = js; j <= je; j++)
        for (int i = is; i <= ie; i++) {
          for (int k = 1; k <= km; k++) dpc[k - 1] = DELP(t, i, j, k);
          a->te0_2d[(size_t)t * ni * nj + (size_t)(j - js) * ni + (i - is)] = column_te(a, t, i, j, dpc, 1, a->pkz);
        }
  double *pe1 = (double *)malloc(sizeof(double) * (ni + 1) * (km + 1));
  double *pe2 = (double *)malloc(sizeof(double) * (ni + 1) * (km + 1));
  double *ws1 = (double *)malloc(sizeof(double) * (ni + 1));
  for (int t = 0; t < nt && !rc; t++) {
    for (int j = js; j <= je + 1 && !rc; j++) {
      /* D-grid u on the y-faces of row j: interfaces averaged over cells j-1, j */
      for (int k = 1; k <= km + 1; k++) {
        const double bkh = 0.5 * a->bk[k - 1];
        for (int i = is; i <= ie; i++) {
          pe1[(size_t)(k - 1) * ni + (i - is)] = 0.5 * (PE(t, i, k, j - 1) + PE(t, i, k, j));
          pe2[(size_t)(k - 1) * ni + (i - is)] = a->ak[k - 1] + bkh * (PS(t, i, j - 1) + PS(t, i, j));
        }
      }
      rc |= fv3o_map1_ppm(km, pe1, a->u + (size_t)t * km * su, NULL, km, pe2, a->u + (size_t)t * km * su, is, ie,
                          j, isd, b->ied, jsd, b->jed + 1, -1, a->kord_mt);
    }
    for (int j = js; j <= je && !rc; j++) {
      /* D-grid v on the x-faces of row j: interfaces averaged over cells i-1, i */
      const int nv = (int)ni + 1;
      for (int k = 1; k <= km + 1; k++) {
        const double bkh = 0.5 * a->bk[k - 1];
        for (int i = is; i <= ie + 1; i++) {
          pe1[(size_t)(k - 1) * nv + (i - is)] = 0.5 * (PE(t, i - 1, k, j) + PE(t, i, k, j));
          pe2[(size_t)(k - 1) * nv + (i - is)] = a->ak[k - 1] + bkh * (PS(t, i - 1, j) + PS(t, i, j));
        }
      }
      rc |= fv3o_map1_ppm(km, pe1, a->v + (size_t)t * km * sv, NULL, km, pe2, a->v + (size_t)t * km * sv, is,
                          ie + 1, j, isd, b->ied + 1, jsd, b->jed, -1, a->kord_mt);
      if (a->hydrostatic) continue;
      /* cell columns: w (bottom value ws) and delz as -delz/delp */
      for (int k = 1; k <= km + 1; k++)
        for (int i = is; i <= ie; i++) {
          pe1[(size_t)(k - 1) * ni + (i - is)] = PE(t, i, k, j);
          pe2[(size_t)(k - 1) * ni + (i - is)] =
              k == 1 ? a->ptop : (k == km + 1 ? PS(t, i, j) : a->ak[k - 1] + a->bk[k - 1] * PS(t, i, j));
        }
      for (int i = is; i <= ie; i++) ws1[i - is] = a->ws[(size_t)t * ni * nj + (size_t)(j - js) * ni + (i - is)];
      rc |= fv3o_map1_ppm(km, pe1, a->w + (size_t)t * km * s2, ws1, km, pe2, a->w + (size_t)t * km * s2, is, ie,
                          j, isd, b->ied, jsd, b->jed, -2, a->kord_wz);
      double *dz = a->delz + (size_t)t * spz;
      for (int k = 1; k <= km; k++)
        for (int i = is; i <= ie; i++)
          dz[(size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)] =
              -dz[(size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)] / DELP(t, i, j, k);
      rc |= fv3o_map1_ppm(km, pe1, dz, NULL, km, pe2, dz, is, ie, j, is, ie, js, je, 1, a->kord_wz);
      for (int k = 1; k <= km; k++)
        for (int i = is; i <= ie; i++)
          dz[(size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)] =
              -dz[(size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)] *
              (pe2[(size_t)k * ni + (i - is)] - pe2[(size_t)(k - 1) * ni + (i - is)]);
    }
  }
  free(pe1);
  free(pe2);
  free(ws1);
  /* tracers, pt, delp <- dp2 */
  for (int t = 0; t < nt && !rc; t++)
    rc |= fv3o_lagrangian_to_eulerian_tracers(km, a->nq, b, a->ak, a->bk, a->ptop, a->delp + (size_t)t * km * s2,
                                              a->q ? a->q + (size_t)t * km * a->nq * s2 : NULL,
                                              a->pt + (size_t)t * km * s2, a->kord_tr, a->kord_tm, a->t_min,
                                              a->fill);
  /* Eulerian pressures: pe, peln, pk, pkz */
  for (int t = 0; t < nt && !rc; t++)
    for (int j = js; j <= je; j++)
      for (int i = is; i <= ie; i++) {
        for (int k = 1; k <= km + 1; k++) {
          const double p = k == 1 ? a->ptop
                                  : (k == km + 1 ? PS(t, i, j) : a->ak[k - 1] + a->bk[k - 1] * PS(t, i, j));
          PE(t, i, k, j) = p;
          const double lp = log(p);
          a->peln[(size_t)t * spl + ((size_t)(j - js) * (km + 1) + (k - 1)) * ni + (i - is)] = lp;
          a->pk[(size_t)t * spk + (size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)] = exp(a->akap * lp);
        }
        for (int k = 1; k <= km; k++) {
          const size_t o = (size_t)t * spz + (size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is);
          if (a->hydrostatic) {
            const double *pkc = a->pk + (size_t)t * spk + (size_t)(j - js) * ni + (i - is);
            const double *plc = a->peln + (size_t)t * spl + (size_t)(j - js) * (km + 1) * ni + (i - is);
            a->pkz[o] = (pkc[(size_t)k * ni * nj] - pkc[(size_t)(k - 1) * ni * nj]) /
                        (a->akap * (plc[(size_t)k * ni] - plc[(size_t)(k - 1) * ni]));
          } else {
            /* p = rho R pt p^akap, so p^akap = (rho R pt)^(akap / (1 - akap)) */
            const double rrg = -a->rdgas / a->grav, k1k = a->akap / (1.0 - a->akap);
            a->pkz[o] = exp(k1k * log(rrg * DELP(t, i, j, k) / a->delz[o] *
                                          a->pt[((size_t)t * km + (k - 1)) * s2 + (size_t)(j - jsd) * nid + (i - isd)]));
          }
        }
      }
  /* total-energy fix-up over the whole domain (compensated sums) */
  if (!rc && a->te0_2d && a->consv > 0.0) {
    ksum e0 = {0, 0}, e1 = {0, 0}, m = {0, 0};
    for (int t = 0; t < nt; t++)
      for (int j = js; j <= je; j++)
        for (int i = is; i <= ie; i++) {
          const double ar = a->area[(size_t)t * s2 + (size_t)(j - jsd) * nid + (i - isd)];
          kadd(&e0, a->te0_2d[(size_t)t * ni * nj + (size_t)(j - js) * ni + (i - is)] * ar);
          kadd(&e1, column_te(a, t, i, j, &DELP(t, i, j, 1), s2, a->pkz) * ar);
          kadd(&m, a->cp * (PS(t, i, j) - a->ptop) * ar);
        }
    const double dtmp = a->consv * (kval(&e0) - kval(&e1)) / kval(&m);
    for (int t = 0; t < nt; t++)
      for (int k = 1; k <= km; k++)
        for (int j = js; j <= je; j++)
          for (int i = is; i <= ie; i++) {
            double *p = &a->pt[((size_t)t * km + (k - 1)) * s2 + (size_t)(j - jsd) * nid + (i - isd)];
            *p = *p + dtmp / a->pkz[(size_t)t * spz + (size_t)(k - 1) * ni * nj + (size_t)(j - js) * ni + (i - is)];
          }
    if (a->dtmp_out) *a->dtmp_out = dtmp;
  }
  free(dpc);
#undef PE
#undef DELP
#undef PS
  return rc;
}
