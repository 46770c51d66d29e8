/*
 * fv3atm_amd.h -- C-ABI of the MI355X tracer transport + vertical remap path
 * (libfv3atm_amd.so).  Plain pointers and sizes only; every function returns
 * 0 on success and a nonzero status on failure, with the reason in
 * fv3_last_error().  Host arrays use the dycore's Fortran layout: column-major,
 * i fastest, explicit lower bounds (isd:ied, jsd:jed, 1:npz, 1:nq); a context
 * created with ntiles = 6 takes six such arrays stacked tile-outermost.
 *
 * The interfaces replaced live in the GFDL_atmos_cubed_sphere submodule, which
 * is empty in the reference (.gitmodules:1-4); they are reached from
 * atmosphere_dynamics (atmos_model.F90:879) via fv_dynamics, and fv_mapz_mod is
 * imported at moving_nest/fv_moving_nest.F90:85.  Argument order and meaning
 * follow the upstream Fortran subroutines named in each comment.
 */
#ifndef FV3ATM_AMD_H
#define FV3ATM_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fv3_ctx fv3_ctx;

/* fv_grid_bounds_type (moving_nest/fv_moving_nest.F90:177-185) */
typedef struct {
  int is, ie, js, je, isd, ied, jsd, jed;
} fv3_bounds;

/* The fv_grid_type components the path reads (moving_nest/fv_moving_nest.F90:81).
 * Host FP64 arrays of one tile: area, rarea, dxa, dya (isd:ied, jsd:jed);
 * dx, del6_u (isd:ied, jsd:jed+1); dy, del6_v (isd:ied+1, jsd:jed);
 * sin_sg (isd:ied, jsd:jed, 9). */
typedef struct {
  const double *area, *rarea, *dx, *dy, *dxa, *dya, *sin_sg, *del6_u, *del6_v;
  double da_min;
  int sw_corner, se_corner, nw_corner, ne_corner, bounded_domain, grid_type;
} fv3_grid;

/* Precision / form selectors of the device-resident step. */
enum {
  FV3_FORM_PLAIN64 = 0, /* staged kernels, FP64 everywhere (parity form) */
  FV3_FORM_FUSED64 = 1, /* fused transport kernel, FP64 */
  FV3_FORM_FUSED32 = 2  /* fused transport kernel, FP32 bulk, FP64 pressure/thickness */
};

/* Thread-local text of the last failure. */
const char *fv3_last_error(void);

/* Context: domain bounds, levels, tile count (1 or 6), the device-resident
 * grid metrics and scratch.  Replaces the state fv_dynamics keeps in
 * gridstruct/bd/domain (the handles tracer_2d receives). */
int fv3_ctx_create(const fv3_bounds *bd, int npx, int npy, int npz, int ntiles, fv3_ctx **out);
int fv3_ctx_destroy(fv3_ctx *ctx);
/* Upload the metrics of tile (1-based). */
int fv3_ctx_set_grid(fv3_ctx *ctx, int tile, const fv3_grid *g);

/* tp_core.F90 fv_tp_2d(q, crx, cry, npx, npy, hord, fx, fy, xfx, yfx,
 *   gridstruct, bd, ra_x, ra_y, lim_fac [, mfx, mfy, mass, nord, damp_c]).
 * Single-tile context.  Optional arguments are NULL.  q's corner halos are
 * modified as upstream. */
int fv3_fv_tp_2d(fv3_ctx *ctx, double *q, const double *crx, const double *cry, int npx,
                 int npy, int hord, double *fx, double *fy, const double *xfx,
                 const double *yfx, const double *ra_x, const double *ra_y, double lim_fac,
                 const double *mfx, const double *mfy, const double *mass, const int *nord,
                 const double *damp_c);

/* fv_tracer2d.F90 tracer_2d_1L(q, dp1, mfx, mfy, cx, cy, gridstruct, bd,
 *   domain, npx, npy, npz, nq, hord, q_split, dt, id_divg, q_pack, dp1_pack,
 *   nord_tr, trdm, lim_fac): per-level sub-cycling, FP64, by the fused
 *   transport kernels (bit for bit the staged form; the staged form runs
 *   when tracer_2d asks for deln_flux damping or id_divg, or runs on a
 *   context of several faces).  q's cube-corner halo blocks come back as
 *   upstream's copy_corners leaves them.
 * All arrays inout as upstream (cx, cy, mfx, mfy scaled by 1/nsplit(k)).
 * nsplit_out (npz, may be NULL) receives the sub-cycle counts. */
int fv3_tracer_2d_1l(fv3_ctx *ctx, double *q, double *dp1, double *mfx, double *mfy,
                     double *cx, double *cy, int npx, int npy, int npz, int nq, int hord,
                     int q_split, double dt, int id_divg, int nord_tr, double trdm,
                     double lim_fac, int *nsplit_out);
/* fv_tracer2d.F90 tracer_2d(...): one sub-cycle count for all levels
 * (q_split > 0 forces it), optional deln_flux on the first sub-cycle. */
int fv3_tracer_2d(fv3_ctx *ctx, double *q, double *dp1, double *mfx, double *mfy, double *cx,
                  double *cy, int npx, int npy, int npz, int nq, int hord, int q_split,
                  double dt, int id_divg, int nord_tr, double trdm, double lim_fac,
                  int *nsplit_out);

/* The same two routines for the dycore's 32-bit build (kind_dyn = 4: the
 * CMake 32BIT / OVERLOAD_R4 build, ci/CMakeLists.txt:20-22,35,59-63): float
 * arrays, computed by the FP32 fused transport (FV3_FORM_FUSED32 kernels)
 * with the in-place semantics above.  deln_flux damping (trdm > 1e-4) and
 * id_divg are not part of this form and fail.  On a context of several
 * faces, tracer_2d_r4 leaves q's halo strips as they came in (upstream
 * leaves the exchange made before its last sub-cycle there). */
int fv3_tracer_2d_1l_r4(fv3_ctx *ctx, float *q, float *dp1, float *mfx, float *mfy, float *cx,
                        float *cy, int npx, int npy, int npz, int nq, int hord, int q_split,
                        double dt, int id_divg, int nord_tr, double trdm, double lim_fac,
                        int *nsplit_out);
int fv3_tracer_2d_r4(fv3_ctx *ctx, float *q, float *dp1, float *mfx, float *mfy, float *cx,
                     float *cy, int npx, int npy, int npz, int nq, int hord, int q_split, double dt,
                     int id_divg, int nord_tr, double trdm, double lim_fac, int *nsplit_out);

/* fv_mapz.F90 map1_ppm(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend,
 *   jbeg, jend, iv, kord): one row j, columns i1:i2; pe1 (i1:i2, km+1),
 *   pe2 (i1:i2, kn+1), q1 (ibeg:iend, jbeg:jend, km), q2 (..., kn), qs (i1:i2). */
int fv3_map1_ppm(int km, const double *pe1, const double *q1, const double *qs, int kn,
                 const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                 int jbeg, int jend, int iv, int kord);
/* fv_mapz.F90 map_scalar(..., q_min): as map1_ppm with the scalar_profile
 * floor q_min. */
int fv3_map_scalar(int km, const double *pe1, const double *q1, const double *qs, int kn,
                   const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                   int jbeg, int jend, int iv, int kord, double q_min);
/* fv_mapz.F90 mapn_tracer(nq, km, pe1, pe2, q1, dp2, kord, j, i1, i2, isd,
 *   ied, jsd, jed, q_min, fill): all tracers of row j, q1 (isd:ied,jsd:jed,km,nq)
 *   inout, dp2 (i1:i2, km), kord (nq). */
int fv3_mapn_tracer(int nq, int km, const double *pe1, const double *pe2, double *q1,
                    const double *dp2, const int *kord, int j, int i1, int i2, int isd, int ied,
                    int jsd, int jed, double q_min, int fill);
/* fv_mapz.F90 cs_profile(qs, a4, delp, km, i1, i2, iv, kord): a4 (4, i1:i2, km)
 * inout, delp (i1:i2, km), qs (i1:i2). */
int fv3_cs_profile(const double *qs, double *a4, const double *delp, int km, int i1, int i2,
                   int iv, int kord);

/* Tracer/pt subset of fv_mapz.F90 Lagrangian_to_Eulerian over every row of
 * every tile of ctx: pe1 = ptop + cumsum(delp), pe2 = ak + bk*ps, tracers by
 * mapn_tracer (nq > 5) or map1_q2 (kord_tr, iv = 0, optional fillz), pt by
 * map_scalar (iv = 1, kord_tm, floor t_min) when pt != NULL, delp <- dp2.
 * delp, pt (isd:ied, jsd:jed, km); q (isd:ied, jsd:jed, km, nq); ak, bk (km+1). */
int fv3_lagrangian_to_eulerian_tracers(fv3_ctx *ctx, int km, int nq, const double *ak,
                                       const double *bk, double ptop, double *delp, double *q,
                                       double *pt, const int *kord_tr, int kord_tm,
                                       double t_min, int fill);

/* The tracer/pt subset above for the 32-bit build: float delp, q, pt (ak, bk
 * double), by the FP32 column-shared remap (FP64 pressures); one kord_tr for
 * every tracer, km >= 8, no fillz. */
int fv3_lagrangian_to_eulerian_tracers_r4(fv3_ctx *ctx, int km, int nq, const double *ak,
                                          const double *bk, double ptop, float *delp, float *q,
                                          float *pt, const int *kord_tr, int kord_tm,
                                          double t_min, int fill);

/* Device-resident advection + remap step (no counterpart upstream as one call:
 * it is tracer_2d_1L followed by the tracer/pt subset of
 * Lagrangian_to_Eulerian, the order fv_dynamics runs them in).  Device
 * pointers; q_in and q_out are distinct buffers of (isd:ied,jsd:jed,npz,nq)
 * per tile.  Winds and dp1 are read-only here: the 1/nsplit scaling is
 * applied in flight.  Element types: form PLAIN64/FUSED64 -> q, pt, cx, cy,
 * mfx, mfy double; FUSED32 -> float.  dp1 (the Lagrangian thickness) and
 * delp_out are always double.  pt (may be NULL) is remapped in place; delp_out
 * receives the target thicknesses.  stream is a hipStream_t (NULL: the
 * context's stream).  nsplit_cap bounds the per-level sub-cycle count the
 * fused forms schedule (0 = derive it, with one device-to-host read). */
int fv3_step_dev(fv3_ctx *ctx, int form, int nq, const void *q_in, void *q_out, void *pt,
                 const double *dp1, double *delp_out, const void *cx, const void *cy,
                 const void *mfx, const void *mfy, const double *ak, const double *bk,
                 double ptop, int hord, int kord_tr, int kord_tm, double t_min,
                 int nsplit_cap, void *stream);

/* Instrumentation (no counterpart upstream, where the "FV3 Dycore" clock of
 * atmos_model.F90:845-847 times the whole dycore): when enabled, each fused
 * fv3_step_dev records hipEvents on its stream around the transport passes
 * and around the remap; fv3_step_timing returns the last step's device times
 * and the number of transport passes launched. */
int fv3_ctx_set_timing(fv3_ctx *ctx, int on);
/* Kernel choices of the FP32 step (no counterpart upstream): cfg selects
 * the fused transport launch shape, waves per workgroup x tracers per wave
 * (0..14; the default is 12: 8 waves x 2 tracers for the interior kernel, 4 x 1
 * at three workgroups per CU for the edge kernels); rows the interior rows per
 * block (0 = 60);
 * remap the vertical remap kernel: 0 (default) = column-shared, a column's
 * levels spread over 16 or 32 lanes of a wave (npz <= 127; larger npz take
 * kernel 2), 1 = one thread per column, 2 = column-shared, one thread per
 * (column, field).  Transport results do not depend on cfg/rows; the remap
 * kernels agree to FP32 rounding. */
int fv3_ctx_set_launch(fv3_ctx *ctx, int cfg, int rows, int remap);
/* No counterpart upstream: run the fused transport's two edge kernels (the
 * blocks that reach a halo or cube-edge row or column) beside the lean
 * interior kernel on auxiliary streams of the context, forked from and
 * joined to the step's stream every pass (on = 1, the default), or after it
 * on the step's stream (on = 0).  Results are identical either way. */
int fv3_ctx_set_concurrency(fv3_ctx *ctx, int on);
/* No counterpart upstream: the FP32 step transports the tracers in chunks
 * of tracers_per_chunk (-1: one round of the launch shape, 8 tracers for the
 * default shape; 0, the default: all tracers at once) and remaps each chunk
 * on an auxiliary stream while the next chunk is transported (with
 * concurrency on).  The per-tracer results do not depend on the chunking. */
int fv3_ctx_set_pipeline(fv3_ctx *ctx, int tracers_per_chunk);
/* Sub-cycling rule of fv3_step_dev (no counterpart upstream; upstream the
 * choice is which routine fv_dynamics calls): per_level = 1 (default) is
 * fv_tracer2d.F90 tracer_2d_1L, nsplit(k) = int(1 + cmax(k)) per level;
 * per_level = 0 is fv_tracer2d.F90 tracer_2d, one nsplit =
 * int(1 + max_k cmax(k)) for every level, or q_split when q_split > 0, with
 * trdm = 0 and id_divg = 0 (the step has no deln_flux damping and does not
 * overwrite dp1 with the divergence). */
int fv3_ctx_set_transport(fv3_ctx *ctx, int per_level, int q_split);
int fv3_step_timing(fv3_ctx *ctx, double *transport_ms, double *remap_ms, int *passes);
/* Graph-safe steps (no counterpart upstream): with on != 0, fv3_step_dev
 * makes no host synchronisation, allocation or memset once its scratch is
 * sized (run one step before capturing), so a caller can capture it in a
 * hipGraph.  It then requires nsplit_cap > 0, skips the phase timing, and
 * leaves the "a level needs more than nsplit_cap sub-cycles" check on the
 * device: fv3_step_status returns it for the last step (1 = overflow, whose
 * results are not valid). */
int fv3_ctx_set_graph_safe(fv3_ctx *ctx, int on);
int fv3_step_status(fv3_ctx *ctx, int *nsplit_overflow);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under hipGraph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The same step as fv3_step_dev in the FP32 form
 * (form must be FV3_FORM_FUSED32, the context must use the column-shared
 * remap, npz >= 8), with every scratch buffer (Courant maxima, sub-cycle
 * counts, the alternate pass buffer, the pt source copy, the remap's column
 * data) carved from one file-scope device array of FV3_NOALLOC_SCRATCH_BYTES
 * per device instead of the context's arenas; it calls no hipMalloc,
 * hipFree or allocating library routine.  A step that needs more scratch
 * fails with a message naming the size.  The array is shared by every call
 * on a device and keeps the last call's contents (each step initialises what
 * it reads); calls that overlap on one device race.  Results are those of
 * fv3_step_dev. */
#define FV3_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)
int fv3_step_dev_noalloc(fv3_ctx *ctx, int form, int nq, const void *q_in, void *q_out, void *pt,
                         const double *dp1, double *delp_out, const void *cx, const void *cy,
                         const void *mfx, const void *mfy, const double *ak, const double *bk,
                         double ptop, int hord, int kord_tr, int kord_tm, double t_min,
                         int nsplit_cap, void *stream);

/* Halo service of a six-tile context: the one-device equivalent of FMS
 * mpp_update_domains on the cubed sphere (upstream fv_dynamics calls it for
 * the prognostic fields; moving_nest/fv_moving_nest.F90:355-387 shows their
 * range).  Device arrays, nslab stacked slabs per tile, tiles outermost;
 * elem_bytes 4 (float) or 8 (double).  The 3-cell edge halos are filled from
 * the neighbouring faces; the 3x3 corner blocks, which lie on no face, are
 * left as they are (copy_corners fills them per sweep).  stream NULL: the
 * context's stream.
 * fv3_halo_update: a scalar at cell centres, slabs (isd:ied, jsd:jed).
 * fv3_halo_update_vector: a vector pair whose components are along each
 * face's +i and +j index directions, turned into the receiving face's
 * directions where the neighbour face is rotated (FMS's vector updates):
 *   FV3_AGRID  u, v at cell centres, (isd:ied, jsd:jed) each;
 *   FV3_CGRID  u on x-faces (isd:ied+1, jsd:jed), v on y-faces (isd:ied, jsd:jed+1);
 *   FV3_DGRID  u on y-faces (isd:ied, jsd:jed+1), v on x-faces (isd:ied+1, jsd:jed)
 *              (the dycore's prognostic winds). */
enum { FV3_AGRID = 1, FV3_CGRID = 2, FV3_DGRID = 3 };
int fv3_halo_update(fv3_ctx *ctx, void *field, int nslab, int elem_bytes, void *stream);
int fv3_halo_update_vector(fv3_ctx *ctx, void *u, void *v, int nslab, int grid, int elem_bytes,
                           void *stream);

/* The full Lagrangian_to_Eulerian of fv_mapz.F90 on device arrays (no
 * single upstream call has this argument list: upstream passes ~60 scalars
 * and arrays; this block carries the ones the remap reads and writes).
 * Element type of the field arrays (u, v, w, delz, pt, q, ws): double for
 * form FV3_FORM_PLAIN64, float for FV3_FORM_FUSED32; pressures, delp and
 * energies are double.  Layouts per tile, tiles outermost:
 *   delp, w, pt (isd:ied, jsd:jed, km); q (isd:ied, jsd:jed, km, nq);
 *   u (isd:ied, jsd:jed+1, km) and v (isd:ied+1, jsd:jed, km): D-grid winds;
 *   delz, pkz (is:ie, js:je, km); ws, te0_2d (is:ie, js:je); ps (isd:ied, jsd:jed);
 *   pe (is-1:ie+1, km+1, js-1:je+1); peln (is:ie, km+1, js:je); pk (is:ie, js:je, km+1).
 * delp holds the Lagrangian thickness on the compute domain and its one-cell
 * ring (the winds' interfaces average neighbouring columns; refresh a six-tile
 * delp with fv3_halo_update first); it returns the Eulerian thickness.
 * Remapped: u, v (kord_mt, iv = -1), w (kord_wz, iv = -2 with the surface
 * value ws) and delz (as -delz/delp, kord_wz) when hydrostatic == 0, the
 * tracers (kord_tr) and pt (kord_tm, floor t_min) as
 * fv3_lagrangian_to_eulerian_tracers.  Recomputed on the Eulerian levels: pe
 * (the ring keeps the Lagrangian interfaces), ps, peln = log pe, pk =
 * pe^akap, and pkz (hydrostatic: the layer mean (pk(k+1)-pk(k))/(akap
 * (peln(k+1)-peln(k))); otherwise (-rdgas/grav delp/delz pt)^(akap/(1-akap)), the p^akap
 * of p = rho R pt p^akap).  With
 * te0_2d != NULL it receives the column total energy before the remap,
 * sum_k (cp pt pkz + (ua^2 + va^2)/2) dp (pkz as passed in, ua/va the face
 * averages of u/v); with consv > 0 the domain-wide loss of that energy is
 * added back as pt += dtmp/pkz, dtmp = consv (E0 - E1) / sum cp (ps-ptop)
 * area, from compensated FP64 sums (dtmp_out, a host pointer or NULL).
 * fast_physics (may be NULL) is called once after the pressures and before
 * the energy fix-up, with the Eulerian state (the GFDL_interstitial fields,
 * ccpp/data/CCPP_typedefs.F90:434-504; kmp by the pfull rule of :1706-1718
 * at p_ref); a nonzero return fails the call. */
typedef struct {
  int npz, nq, ntiles, kmp, last_step;
  double mdt;
  double *delp, *pe, *peln, *pk, *pkz, *ps, *te0_2d;
  void *pt, *q, *delz, *w;
  void *stream;
} fv3_l2e_state;
typedef int (*fv3_fast_physics_fn)(void *user, const fv3_l2e_state *state);
typedef struct {
  int km, nq;
  const double *ak, *bk;
  double ptop, akap, cp, rdgas, grav, consv, t_min, p_ref, mdt;
  int hydrostatic, fill, kord_mt, kord_wz, kord_tm, last_step;
  const int *kord_tr; /* host array (nq) */
  double *delp;
  void *u, *v, *w, *delz, *pt, *q;
  const void *ws;
  double *pe, *peln, *pk, *pkz, *ps, *te0_2d, *dtmp_out;
  fv3_fast_physics_fn fast_physics;
  void *fast_physics_user;
} fv3_l2e_args;
int fv3_lagrangian_to_eulerian_dev(fv3_ctx *ctx, int form, const fv3_l2e_args *args, void *stream);
/* GFDL_interstitial's kmp (ccpp/data/CCPP_typedefs.F90:1706-1718): the first
 * level whose full-level pressure at p_ref exceeds 10 hPa (host arrays). */
int fv3_kmp(int npz, double p_ref, const double *ak, const double *bk, int *kmp);

/* The tracer path's inputs as dyn_core accumulates them over its n_split
 * acoustic sub-steps (fv_dynamics hands them to tracer_2d_1L; dyn_core and
 * d_sw are in the absent atmos_cubed_sphere submodule, reached from
 * atmos_model.F90:879).  Each sub-step of length bdt turns the sub-step's
 * C-grid winds into Courant numbers crx = bdt uc / dxa(upwind) and area
 * fluxes xfx = crx dxa dy sin_sg(upwind) (cry, yfx likewise), transports
 * delp with fv_tp_2d (order hord_dp, fluxes = area flux x delp face value,
 * copy_corners acting on delp), adds crx, cry and the delp fluxes into cx,
 * cy, mfx, mfy, updates delp by the flux divergence and, on six faces,
 * refreshes delp's edge halos before the next sub-step (one tile: frozen).
 * FV3_FORM_PLAIN64 runs each sub-step as a sequence of stage kernels;
 * FV3_FORM_FUSED64 and FV3_FORM_FUSED32 run it as one launch whose blocks keep
 * the intermediates on chip (both FP64 forms equal the FP64 oracle bit for
 * bit).  Device arrays, tiles outermost; element type double
 * (FV3_FORM_PLAIN64, FV3_FORM_FUSED64) or float (FV3_FORM_FUSED32):
 *   uc: n_split sub-steps of (is:ie+1, jsd:jed, npz) per tile (sub-step
 *       outermost), vc: of (isd:ied, js:je+1, npz);
 *   delp (isd:ied, jsd:jed, npz): in, the thickness at the start with its
 *       halo filled; out, the thickness after the sub-steps;
 *   dp1: out, delp at the start; cx (is:ie+1, jsd:jed, npz), cy (isd:ied,
 *       js:je+1, npz), mfx (is:ie+1, js:je, npz), mfy (is:ie, js:je+1, npz):
 *       out, the sums over the sub-steps.
 * dp1 + (mfx(i) - mfx(i+1) + mfy(j) - mfy(j+1)) rarea is the final delp up
 * to rounding: the consistency that keeps tracer_2d's constants constant. */
int fv3_dyn_core_fluxes(fv3_ctx *ctx, int form, int n_split, double bdt, int hord_dp, const void *uc,
                        const void *vc, void *delp, void *dp1, void *cx, void *cy, void *mfx, void *mfy,
                        void *stream);

/* Global tracer mass: out[iq] = the sum over every tile, level and
 * compute-domain cell of q * delp * area (the product in that order, FP64),
 * with Neumaier-compensated FP64 sums in a fixed reduction order -- the same
 * result on every run, within a few units in the last place of the exact
 * sum of the products.  It stands for the dycore's global mass sums (g_sum
 * over mpp_global_sum; the reference imports its bit-reproducible mode,
 * moving_nest/fv_moving_nest_utils.F90:35,41).  q (isd:ied, jsd:jed, npz, nq)
 * and delp (isd:ied, jsd:jed, npz) per tile on the device, q float or
 * double; out: host, nq values. */
int fv3_global_mass(fv3_ctx *ctx, const void *q, int nq, const double *delp, int elem_bytes, double *out,
                    void *stream);

/* ---- on-disk formats ----------------------------------------------------------
 * Files are netCDF classic with 64-bit offsets (CDF-2: the NF90_64BIT_OFFSET
 * choice of io/module_write_netcdf.F90:35 and FMS's restart default), written
 * without libnetcdf.  Field arrays are device arrays with the halo'd layout
 * (isd:ied, jsd:jed, nslab) per tile, tiles outermost; files hold the compute
 * domain.  stream NULL: the context's stream.
 *
 * fv3_field_checksum: FMS mpp_chksum of each tile's compute domain -- the sum
 * modulo 2^64 of the values' bit patterns read as two's-complement integers
 * (float values: their 32-bit patterns, sign-extended).  out: ntiles values. */
int fv3_field_checksum(fv3_ctx *ctx, const void *field, int nslab, int elem_bytes,
                       unsigned long long *out, void *stream);

/* The dycore's tracer restart, one file per tile named <prefix>.tile<t>.nc
 * for t = tile_base .. tile_base + ntiles - 1 (upstream fv_io writes
 * fv_tracer.res.tile<t>.nc beside the physics restarts of
 * io/fv3atm_restart_io.F90:126-146): dimensions xaxis_1, yaxis_1, zaxis_1 and
 * the record dimension Time; one double variable (Time, zaxis_1, yaxis_1,
 * xaxis_1) per tracer named names[iq], with long_name, units and the FMS
 * checksum attribute (16 upper-case hex digits, right-justified).
 * q: (isd:ied, jsd:jed, npz, nq) per tile, float or double.  The read fills
 * the compute domain of q (halos untouched) from the last record and, unless
 * ignore_cksum (atmos_model.F90:193-194), fails when a tracer's checksum
 * differs from its attribute. */
int fv3_tracer_restart_write(fv3_ctx *ctx, const char *prefix, int tile_base, int nq,
                             const char *const *names, const void *q, int elem_bytes, void *stream);
int fv3_tracer_restart_read(fv3_ctx *ctx, const char *prefix, int tile_base, int nq,
                            const char *const *names, void *q, int elem_bytes, int ignore_cksum,
                            void *stream);

/* The dycore history bundle in the layout of the write component's netCDF
 * writer (io/module_write_netcdf.F90:50-560, cubed-sphere output grid):
 * dimensions grid_xt, grid_yt, nchars, pfull, phalf, tile, time (fixed, 1);
 * variables pfull, phalf (float), time, grid_xt, grid_yt (1..n), tile,
 * time_iso, lon, lat (degrees, host (is:ie, js:je) per tile; NULL: zeros),
 * then each field as (time, tile, [pfull,] grid_yt, grid_xt) of out_bytes
 * (4: NF90_FLOAT, 8: NF90_DOUBLE) with long_name, units, coordinates and
 * grid_mapping.  quantize_nsd > 0 applies BitRound to the 3-D fields
 * (module_write_netcdf.F90:417-443, quantize_mode 'quantize_bitround': keep
 * nsd mantissa bits, a half rounds away from zero) and records
 * _QuantizeBitRoundNumberOfSignificantBits.  Global attributes: grid, im,
 * jm and, when given, ak and bk (npz + 1 each). */
typedef struct {
  const char *name, *long_name, *units;
  const void *data; /* device, (isd:ied, jsd:jed, nlev) per tile */
  int nlev;         /* 1 (a 2-D field) or npz */
  int elem_bytes;   /* 4 or 8 */
} fv3_history_field;
typedef struct {
  int nfields;
  const fv3_history_field *fields;
  int out_bytes, quantize_nsd, tile_base;
  const float *pfull, *phalf; /* host, npz and npz + 1 */
  const double *lon, *lat, *ak, *bk;
  double time;
  const char *time_units, *time_iso;
} fv3_history_args;
int fv3_history_write(fv3_ctx *ctx, const char *path, const fv3_history_args *args, void *stream);

/* Tangent-linear and adjoint of the FP64 advection + tracer remap step: the
 * derivatives of tracer_2d (per_level = 0; q_split > 0 forces nsplit) or
 * tracer_2d_1L (per_level = 1) followed by the tracer part of
 * Lagrangian_to_Eulerian (kord_tr; 0 leaves the remap out), with respect to
 * the interior tracer values, every limiter branch, upwind choice and
 * source-layer search frozen at the forward trajectory q0.  Winds and dp1
 * are fixed; a single tile's halo is fixed too, while on six faces every
 * sub-cycle refreshes the halos (values and tangents) from the neighbouring
 * faces and the adjoint carries the halo adjoints back onto the cells they
 * were copied from.  Device pointers, layouts as
 * fv3_step_dev, all double.  The upstream dycore has no such routines; they
 * serve variational and sensitivity work on this path.
 * TL: (q0, dq0) -> (q1, dq1); q1 is the forward result, delp_out (may be
 * NULL) the target thickness.  AD: qbar1 (adjoint of q1) -> qbar0 (adjoint of
 * q0's interior; halo cells 0).  The adjoint recomputes each sub-cycle from
 * checkpointed states. */
int fv3_step_tl_dev(fv3_ctx *ctx, int per_level, int q_split, int nq, const double *q0,
                    const double *dq0, double *q1, double *dq1, const double *dp1,
                    double *delp_out, const double *cx, const double *cy, const double *mfx,
                    const double *mfy, const double *ak, const double *bk, double ptop, int hord,
                    int kord_tr, void *stream);
int fv3_step_ad_dev(fv3_ctx *ctx, int per_level, int q_split, int nq, const double *q0,
                    const double *qbar1, double *qbar0, const double *dp1, const double *cx,
                    const double *cy, const double *mfx, const double *mfy, const double *ak,
                    const double *bk, double ptop, int hord, int kord_tr, void *stream);

#ifdef __cplusplus
}
#endif
#endif
