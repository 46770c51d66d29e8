/*
 * fv3_oracle.h -- CPU FP64 restatement of FV3's tracer transport and vertical
 * remap path.  TEST INFRASTRUCTURE ONLY: this library is the checker used by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The
 * product (libfv3atm_amd.so) never links or calls it.
 *
 * The upstream sources (GFDL_atmos_cubed_sphere model/tp_core.F90,
 * model/fv_tracer2d.F90, model/fv_mapz.F90) are an empty submodule in the
 * reference (.gitmodules:1-4); every routine here is restated from the
 * published algorithm (Lin & Rood 1996, Lin 2004, Putman & Lin 2007,
 * Colella & Woodward 1984, Huynh 1996) and from the public upstream interface.
 * Arrays follow the Fortran conventions of the dycore: column-major, i
 * fastest, lower bounds carried explicitly (isd:ied, jsd:jed, 1:npz, 1:nq),
 * as at ccpp/data/CCPP_typedefs.F90:1580-1590 and
 * moving_nest/fv_moving_nest.F90:177-189.
 *
 * Parity status.  The reference holds no golden vectors, known-answer tests
 * or fixtures for this path, and its upstream source is absent, so no
 * function can be compared with upstream output.  Each function is pinned
 * instead by what the mathematics and the geometry fix (tests/test_oracle_*.py,
 * tests/test_golden_rational.py):
 *   xppm/yppm hord 8 (interior and cube edge): exact-rational restatement on
 *     8-cell lines (tests/golden, hashed fixtures); exact on linear data (all
 *     orders) and parabola means; the cube-edge face value exact for data
 *     linear in distance on the real gnomonic spacing; unit Courant takes the
 *     upwind cell; yppm is xppm transposed; hord 8/11 obey the discrete max
 *     principle.
 *   pert_ppm iv=1: the constrained parabola is monotone in the cell.
 *   copy_corners: each corner cell copies the cell that continues the sweep's
 *     grid line around the cube corner on the sphere; dir 1 and dir 2 agree
 *     under the diagonal swap.
 *   fv_tp_2d, tracer_2d_1L, tracer_2d: constant preservation, six-tile mass
 *     conservation to 1e-13, second-order convergence away from the cube
 *     edges (C24-C96), a full cosine-bell revolution, six-tile rotation
 *     symmetry, the nsplit rule against hand-computed counts, forced q_split.
 *   deln_flux: zero flux on constants, variance reduction on noise.
 *   cs_profile kord 9 + map1_ppm: exact-rational restatement on 6- and
 *     10-layer columns (tests/golden); C1 spline when unlimited, exact on
 *     linear profiles; 2-delta-p waves flattened, one-layer extrema kept
 *     curved, smooth monotone profiles unconstrained (kord 9 and 10).
 *   map1_ppm, map_scalar, mapn_tracer, Lagrangian_to_Eulerian subset:
 *     identity remap, column conservation, constancy, refine-then-coarsen,
 *     iv=0 positivity, pressure bookkeeping.
 *   fillz: conservation and non-negativity.
 * Parity unpinned (only the absent upstream source could tell a faithful
 * restatement from a plausible variant): the constants of the one-sided
 * cube-edge 3-point face values (11/14, 3/14, 4/7) and of the cs_profile end
 * closures beyond their linear exactness, the 1 - sin_sg(5) term of the nsplit
 * rule, the hord 9/10/12/13 Huynh-constraint details beyond linear
 * exactness, positivity and the max principle, and the kord 10 extremum
 * branches beyond the wave/extremum rules above.
 */
#ifndef FV3_ORACLE_H
#define FV3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* fv_grid_bounds_type (moving_nest/fv_moving_nest.F90:177-185). */
typedef struct {
  int is, ie, js, je;
  int isd, ied, jsd, jed;
} fv3o_bounds;

/* The fv_grid_type components the path reads (moving_nest/fv_moving_nest.F90:81).
 * Shapes: area, rarea, dxa, dya: (isd:ied, jsd:jed); dx, del6_u: (isd:ied,
 * jsd:jed+1); dy, del6_v: (isd:ied+1, jsd:jed); sin_sg: (isd:ied, jsd:jed, 9). */
typedef struct {
  const double *area, *rarea, *dx, *dy, *dxa, *dya, *sin_sg, *del6_u, *del6_v;
  double da_min;
  int sw_corner, se_corner, nw_corner, ne_corner;
  int bounded_domain, grid_type;
} fv3o_grid;

/* ---- tp_core ---------------------------------------------------------- */
void fv3o_copy_corners(double *q, int npx, int npy, int dir, const fv3o_bounds *bd,
                       const fv3o_grid *g);
void fv3o_pert_ppm(int im, const double *a0, double *al, double *ar, int iv);
/* xppm: q(isd:ied, jfirst:jlast), c(is:ie+1, jfirst:jlast), flux(is:ie+1, jfirst:jlast) */
int fv3o_xppm(double *flux, const double *q, const double *c, int iord, int is, int ie,
              int isd, int ied, int jfirst, int jlast, int jsd, int jed, int npx, int npy,
              const double *dxa, int bounded_domain, int grid_type, double lim_fac);
/* yppm: q(ifirst:ilast, jsd:jed), c(isd:ied, js:je+1), flux(ifirst:ilast, js:je+1) */
int fv3o_yppm(double *flux, const double *q, const double *c, int jord, int ifirst,
              int ilast, int isd, int ied, int js, int je, int jsd, int jed, int npx,
              int npy, const double *dya, int bounded_domain, int grid_type, double lim_fac);
void fv3o_deln_flux(int nord, int npx, int npy, double damp, const double *q, double *fx,
                    double *fy, const fv3o_grid *g, const fv3o_bounds *bd, const double *mass);
/* Optional Fortran arguments are NULL pointers here. */
int fv3o_fv_tp_2d(double *q, const double *crx, const double *cry, int npx, int npy, int hord,
                  double *fx, double *fy, const double *xfx, const double *yfx,
                  const fv3o_grid *g, const fv3o_bounds *bd, const double *ra_x,
                  const double *ra_y, double lim_fac, const double *mfx, const double *mfy,
                  const double *mass, const int *nord, const double *damp_c);

/* ---- fv_tracer2d ------------------------------------------------------- */
/* Halo refresh hook between sub-cycles: called with (ctx, field, nlev-or-nq
 * stacked 2-D slabs, level k (1-based) or 0 when all levels are refreshed).
 * NULL keeps halos as given at entry (single tile, frozen-halo policy). */
typedef void (*fv3o_halo_fn)(void *ctx, double *field, int nslab, int k);

/* ---- dyn_core's tracer inputs (fv3_oracle_dyn.c) ------------------------
 * uc: n_split sub-steps of (is:ie+1, jsd:jed, npz) per tile, vc: of
 * (isd:ied, js:je+1, npz); delp in: start thickness (halo filled), out: the
 * thickness after the sub-steps; dp1 out: delp at entry; cx, cy, mfx, mfy
 * out: the accumulated Courant numbers and delp fluxes. */
int fv3o_dyn_core_fluxes(int n_split, double bdt, int hord_dp, const double *uc, const double *vc,
                         double *delp, double *dp1, double *cx, double *cy, double *mfx, double *mfy,
                         const fv3o_grid *g, const fv3o_bounds *bd, int ntiles, int npx, int npy,
                         int npz, double lim_fac, fv3o_halo_fn halo, void *halo_ctx);

/* Per-level sub-cycling (tracer_2d_1L).  ntiles tiles are stacked outermost
 * in every array (tile stride = the single-tile array size).  nsplit_out
 * (npz) receives nsplit(k) when non-NULL. */
int fv3o_tracer_2d_1L(double *q, double *dp1, double *mfx, double *mfy, double *cx,
                      double *cy, const fv3o_grid *g, const fv3o_bounds *bd, int ntiles,
                      int npx, int npy, int npz, int nq, int hord, int q_split, double dt,
                      int id_divg, int nord_tr, double trdm, double lim_fac,
                      fv3o_halo_fn halo, void *halo_ctx, int *nsplit_out);
/* Global sub-cycling (tracer_2d). */
int fv3o_tracer_2d(double *q, double *dp1, double *mfx, double *mfy, double *cx, double *cy,
                   const fv3o_grid *g, const fv3o_bounds *bd, int ntiles, int npx, int npy,
                   int npz, int nq, int hord, int q_split, double dt, int id_divg,
                   int nord_tr, double trdm, double lim_fac, fv3o_halo_fn halo,
                   void *halo_ctx, int *nsplit_out);

/* ---- fv_mapz ------------------------------------------------------------ */
/* a4(4, i1:i2, km), delp(i1:i2, km), qs(i1:i2) */
void fv3o_cs_limiters(int im, const int *extm, double *a4, int iv);
int fv3o_cs_profile(const double *qs, double *a4, const double *delp, int km, int i1, int i2,
                    int iv, int kord);
int fv3o_scalar_profile(const double *qs, double *a4, const double *delp, int km, int i1,
                        int i2, int iv, int kord, double qmin);
int fv3o_map1_ppm(int km, const double *pe1, const double *q1, const double *qs, int kn,
                  const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                  int jbeg, int jend, int iv, int kord);
int fv3o_map_scalar(int km, const double *pe1, const double *q1, const double *qs, int kn,
                    const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                    int jbeg, int jend, int iv, int kord, double q_min);
int fv3o_map1_q2(int km, const double *pe1, const double *q1, int kn, const double *pe2,
                 double *q2, const double *dp2, int i1, int i2, int iv, int kord, int j,
                 int ibeg, int iend, int jbeg, int jend, double q_min);
int fv3o_mapn_tracer(int nq, int km, const double *pe1, const double *pe2, double *q1,
                     const double *dp2, const int *kord, int j, int i1, int i2, int isd,
                     int ied, int jsd, int jed, double q_min, int fill);
void fv3o_fillz(int im, int km, int nq, double *q, const double *dp);
/* Tracer/pt subset of Lagrangian_to_Eulerian over all rows js:je.
 * delp, pt (isd:ied, jsd:jed, km); q (isd:ied, jsd:jed, km, nq); ak, bk (km+1).
 * pt is remapped with map_scalar (iv=1, kord_tm, t_min) when pt != NULL. */
int fv3o_lagrangian_to_eulerian_tracers(int km, int nq, const fv3o_bounds *bd,
                                        const double *ak, const double *bk, double ptop,
                                        double *delp, double *q, double *pt,
                                        const int *kord_tr, int kord_tm, double t_min,
                                        int fill);

/* The rest of Lagrangian_to_Eulerian (fv3_oracle_l2e.c).  Arrays per tile,
 * tiles outermost: delp, w, pt (isd:ied, jsd:jed, km); q (..., km, nq);
 * u (isd:ied, jsd:jed+1, km) and v (isd:ied+1, jsd:jed, km) on the D grid;
 * delz, pkz (is:ie, js:je, km); ws, te0_2d (is:ie, js:je); ps, area
 * (isd:ied, jsd:jed); pe (is-1:ie+1, km+1, js-1:je+1); peln (is:ie, km+1,
 * js:je); pk (is:ie, js:je, km+1).  delp must hold the Lagrangian thickness
 * on the compute domain and its one-cell ring (the winds' interfaces average
 * neighbouring columns).  w and delz are remapped when hydrostatic is 0;
 * te0_2d (may be NULL) receives the column energy before the remap, and with
 * consv > 0 the domain's energy loss is added back to pt (dtmp_out, may be
 * NULL, receives the increment). */
typedef struct {
  int km, nq, ntiles;
  fv3o_bounds bd;
  const double *ak, *bk;
  double ptop, akap, cp, rdgas, grav, consv, t_min;
  int hydrostatic, fill, kord_mt, kord_wz, kord_tm;
  const int *kord_tr;
  double *delp, *u, *v, *w, *delz, *pt, *q;
  const double *ws, *area;
  double *pe, *peln, *pk, *pkz, *ps, *te0_2d, *dtmp_out;
} fv3o_l2e;
int fv3o_lagrangian_to_eulerian_full(const fv3o_l2e *a);

/* OpenMP threads of the level (tracer_2d_1L) and row (remap) loops; n <= 0
 * keeps the current count.  Returns the count in effect (1 without OpenMP). */
int fv3o_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
