/* fv3_oracle_dyn.c -- FP64 CPU restatement of what dyn_core hands the tracer
 * path: the accumulated Courant numbers cx, cy, mass fluxes mfx, mfy and the
 * start thickness dp1 (fv_dynamics passes them to tracer_2d_1L after the
 * acoustic loop; dyn_core and d_sw live in the absent atmos_cubed_sphere
 * submodule, .gitmodules:1-4, reached from atmos_model.F90:879).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the GPU path).
 *
 * The published scheme (Lin 2004, "A vertically Lagrangian finite-volume
 * dynamical core for global models", sect. 3; the FV3 scientific
 * description, Harris et al. 2021, sect. 4-5): each of n_split acoustic
 * sub-steps of length bdt
 *   - turns the sub-step's C-grid winds into Courant numbers with the upwind
 *     cell width, crx = bdt uc / dxa(upwind), and area fluxes through the
 *     face with the upwind metric, xfx = crx dxa dy sin_sg(upwind) (the
 *     relation tracer_2d itself uses between cx and xfx);
 *   - transports delp with fv_tp_2d (order hord_dp, no mass fluxes: the
 *     fluxes are xfx times the delp face values), copy_corners acting on
 *     delp in place;
 *   - adds the delp fluxes to mfx, mfy and the Courant numbers to cx, cy;
 *   - updates delp by the flux divergence,
 *     delp += ((fx(i) - fx(i+1)) + (fy(j) - fy(j+1))) rarea;
 *   - refreshes delp's halo before the next sub-step (six tiles: the halo
 *     callback; one tile: halos frozen).
 * dp1 is delp at entry.  By construction dp1 + div(mfx, mfy) rarea is the
 * final delp up to rounding, which is what keeps tracer_2d's constant
 * tracers constant.
 *
 * Parity unpinned: the exact upstream operation order of crx and xfx (the
 * source is absent); the tests pin the consistency relations instead. */
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include "fv3_oracle.h"

#define E2(p, i, j, il, jl, ni) (p)[(ptrdiff_t)((j) - (jl)) * (ni) + ((i) - (il))]

int fv3o_dyn_core_fluxes(int n_split, double bdt, int hord_dp, const double *uc, const double *vc,
                         double *delp, double *dp1, double *cx, double *cy, double *mfx, double *mfy,
                         const fv3o_grid *g, const fv3o_bounds *bd, int ntiles, int npx, int npy,
                         int npz, double lim_fac, fv3o_halo_fn halo, void *halo_ctx) {
  const int is = bd->is, ie = bd->ie, js = bd->js, je = bd->je;
  const int isd = bd->isd, ied = bd->ied, jsd = bd->jsd, jed = bd->jed;
  const int nid = ied - isd + 1, njd = jed - jsd + 1, nix = ie - is + 2, niy = ie - is + 1;
  const size_t s2 = (size_t)nid * njd, sx = (size_t)nix * njd, sy = (size_t)nid * (je - js + 2);
  const size_t smx = (size_t)nix * (je - js + 1), smy = (size_t)niy * (je - js + 2);
  const size_t nz = (size_t)ntiles * npz;
  int rc = 0;
  double *crx = malloc(sizeof(double) * sx), *xfx = malloc(sizeof(double) * sx);
  double *cry = malloc(sizeof(double) * sy), *yfx = malloc(sizeof(double) * sy);
  double *rax = malloc(sizeof(double) * (size_t)niy * njd);
  double *ray = malloc(sizeof(double) * (size_t)nid * (je - js + 1));
  double *fx = malloc(sizeof(double) * smx), *fy = malloc(sizeof(double) * smy);
  memcpy(dp1, delp, sizeof(double) * s2 * nz);
  memset(cx, 0, sizeof(double) * sx * nz);
  memset(cy, 0, sizeof(double) * sy * nz);
  memset(mfx, 0, sizeof(double) * smx * nz);
  memset(mfy, 0, sizeof(double) * smy * nz);
  for (int it = 0; it < n_split && !rc; it++) {
    for (int t = 0; t < ntiles && !rc; t++) {
      const fv3o_grid *gt = &g[t];
      for (int k = 0; k < npz && !rc; k++) {
        const size_t z = (size_t)t * npz + k;
        const double *u = uc + ((size_t)it * nz + z) * sx, *v = vc + ((size_t)it * nz + z) * sy;
        for (int j = jsd; j <= jed; j++)
          for (int i = is; i <= ie + 1; i++) {
            const double w = E2(u, i, j, is, jsd, nix);
            const int iu = w > 0.0 ? i - 1 : i;
            const double c = bdt * w / E2(gt->dxa, iu, j, isd, jsd, nid);
            E2(crx, i, j, is, jsd, nix) = c;
            E2(xfx, i, j, is, jsd, nix) = c * E2(gt->dxa, iu, j, isd, jsd, nid) *
                                          E2(gt->dy, i, j, isd, jsd, nid + 1) *
                                          E2(gt->sin_sg + (w > 0.0 ? 2 : 0) * s2, iu, j, isd, jsd, nid);
          }
        for (int j = js; j <= je + 1; j++)
          for (int i = isd; i <= ied; i++) {
            const double w = E2(v, i, j, isd, js, nid);
            const int ju = w > 0.0 ? j - 1 : j;
            const double c = bdt * w / E2(gt->dya, i, ju, isd, jsd, nid);
            E2(cry, i, j, isd, js, nid) = c;
            E2(yfx, i, j, isd, js, nid) = c * E2(gt->dya, i, ju, isd, jsd, nid) *
                                          E2(gt->dx, i, j, isd, jsd, nid) *
                                          E2(gt->sin_sg + (w > 0.0 ? 3 : 1) * s2, i, ju, isd, jsd, nid);
          }
        for (int j = jsd; j <= jed; j++)
          for (int i = is; i <= ie; i++)
            E2(rax, i, j, is, jsd, niy) = E2(gt->area, i, j, isd, jsd, nid) + E2(xfx, i, j, is, jsd, nix) -
                                          E2(xfx, i + 1, j, is, jsd, nix);
        for (int j = js; j <= je; j++)
          for (int i = isd; i <= ied; i++)
            E2(ray, i, j, isd, js, nid) = E2(gt->area, i, j, isd, jsd, nid) + E2(yfx, i, j, isd, js, nid) -
                                          E2(yfx, i, j + 1, isd, js, nid);
        double *dk = delp + z * s2;
        rc |= fv3o_fv_tp_2d(dk, crx, cry, npx, npy, hord_dp, fx, fy, xfx, yfx, gt, bd, rax, ray, lim_fac,
                            NULL, NULL, NULL, NULL, NULL);
        double *cxk = cx + z * sx, *cyk = cy + z * sy, *mxk = mfx + z * smx, *myk = mfy + z * smy;
        for (size_t m = 0; m < sx; m++) cxk[m] = cxk[m] + crx[m];
        for (size_t m = 0; m < sy; m++) cyk[m] = cyk[m] + cry[m];
        for (size_t m = 0; m < smx; m++) mxk[m] = mxk[m] + fx[m];
        for (size_t m = 0; m < smy; m++) myk[m] = myk[m] + fy[m];
        for (int j = js; j <= je; j++)
          for (int i = is; i <= ie; i++)
            E2(dk, i, j, isd, jsd, nid) =
                E2(dk, i, j, isd, jsd, nid) +
                ((E2(fx, i, j, is, js, nix) - E2(fx, i + 1, j, is, js, nix)) +
                 (E2(fy, i, j, is, js, niy) - E2(fy, i, j + 1, is, js, niy))) *
                    E2(gt->rarea, i, j, isd, jsd, nid);
      }
    }
    if (halo && it + 1 < n_split) halo(halo_ctx, delp, npz, 0);
  }
  free(crx);
  free(xfx);
  free(cry);
  free(yfx);
  free(rax);
  free(ray);
  free(fx);
  free(fy);
  return rc;
}
