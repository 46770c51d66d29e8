// dyncore.hip -- the tracer path's inputs as dyn_core accumulates them over
// its acoustic sub-steps: the Courant numbers cx, cy, the mass fluxes mfx,
// mfy and the start thickness dp1 that fv_dynamics hands to tracer_2d_1L
// (dyn_core / d_sw live in the absent atmos_cubed_sphere submodule, reached
// from atmos_model.F90:879).  Each of n_split sub-steps of length bdt
//   - turns the sub-step's C-grid winds into Courant numbers with the upwind
//     cell width and area fluxes through the face with the upwind metric,
//   - transports delp with fv_tp_2d (no mass fluxes: the fluxes are the
//     area fluxes times the delp face values; copy_corners acts on delp),
//   - adds the delp fluxes and Courant numbers into mfx, mfy, cx, cy,
//   - updates delp by the flux divergence, and
//   - refreshes delp's six-face halo before the next sub-step.
// Every (tile, level) slab of a sub-step runs in one launch per stage; the
// stage kernels are those of the staged transport, in the FP64 oracle's
// operation order, so the FP64 form matches it bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "../../include/fv3atm_amd.h"
#include "fv3_internal.h"

namespace fv3 {

namespace {

template <typename F>
__global__ void __launch_bounds__(256) k_rect(F f, int i0, int i1, int j0, int j1) {
  const int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i <= i1 && j <= j1) f(i, j, (int)blockIdx.z);
}

template <typename F>
int launch_rect(hipStream_t s, F f, int i0, int i1, int j0, int j1, int nz) {
  if (i1 < i0 || j1 < j0 || nz <= 0) return 0;
  const dim3 b(64, 4);
  const dim3 g((unsigned)((i1 - i0 + 64) / 64), (unsigned)((j1 - j0 + 4) / 4), (unsigned)nz);
  hipLaunchKernelGGL(k_rect<F>, g, b, 0, s, f, i0, i1, j0, j1);
  FV3_HIP(hipGetLastError());
  return 0;
}

template <typename T>
struct RowQ {
  const T *p;
  int lo;
  __device__ __forceinline__ T operator()(int m) const { return p[m - lo]; }
};
template <typename T>
struct ColQ {
  const T *p;
  int lo;
  long long stride;
  __device__ __forceinline__ T operator()(int m) const { return p[(long long)(m - lo) * stride]; }
};

}  // namespace

template <typename T>
static int dyn_core_fluxes(Ctx &c, int n_split, T bdt, int hord, const T *uc, const T *vc, T *delp, T *dp1,
                           T *cx, T *cy, T *mfx, T *mfy, hipStream_t s) {
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles, nz = npz * nt;
  const GridDev<T> g = grid_of<T>(c);
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  const long long sz = nz;
  size_t off = 0;
  auto take = [&](long long n) {
    const size_t o = off;
    off += ((size_t)n * sizeof(T) + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_crx = take(d.sx * sz), o_xfx = take(d.sx * sz), o_cry = take(d.sy * sz), o_yfx = take(d.sy * sz);
  const size_t o_rax = take(d.srx * sz), o_ray = take(d.sry * sz), o_fx = take(d.smx * sz), o_fy = take(d.smy * sz);
  const size_t o_fy2 = take(d.sy * sz), o_qi = take(d.sry * sz), o_fx2 = take(d.sx * sz), o_qj = take(d.srx * sz);
  char *base = (char *)ctx_scratch(c, off);
  if (!base) return 1;
  T *crx = (T *)(base + o_crx), *xfx = (T *)(base + o_xfx), *cry = (T *)(base + o_cry), *yfx = (T *)(base + o_yfx);
  T *rax = (T *)(base + o_rax), *ray = (T *)(base + o_ray), *fx = (T *)(base + o_fx), *fy = (T *)(base + o_fy);
  T *t_fy2 = (T *)(base + o_fy2), *t_qi = (T *)(base + o_qi), *t_fx2 = (T *)(base + o_fx2), *t_qj = (T *)(base + o_qj);

  FV3_HIP(hipMemcpyAsync(dp1, delp, sizeof(T) * d.s2 * sz, hipMemcpyDeviceToDevice, s));
  const long long gs2 = g.s2, gdx = g.sdx, gdy = g.sdy;
  const T *dxa = g.dxa, *dya = g.dya, *gdxp = g.dx, *gdyp = g.dy, *sg = g.sin_sg, *area = g.area,
          *rarea = g.rarea;
  const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
  const unsigned corners = d.corners;
  auto cc = [&](int dir) {
    if (d.bounded) return 0;
    return launch_rect(s, [=] __device__(int cidx, int, int z) {
          T *w = delp + (long long)z * d.s2;
          const int blk = cidx / 9, k9 = cidx % 9;
          const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + k9 % 3;
          const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + k9 / 3;
          int si, sj;
          if (corner_src(i, j, d.npx, d.npy, dir, corners, si, sj))
            w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = w[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
        }, 0, 35, 0, 0, nz);
  };

  for (int it = 0; it < n_split; it++) {
    const T *u = uc + (long long)it * sz * d.sx, *v = vc + (long long)it * sz * d.sy;
    const bool first = it == 0;
    // Courant numbers and area fluxes with the upwind metric; cx, cy accumulate
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const long long lv = (long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
          const T w = u[lv];
          const int iu = w > T(0) ? i - 1 : i;
          const long long m = tile * gs2 + (long long)(j - d.jsd) * d.nid + (iu - d.isd);
          const T cr = bdt * w / dxa[m];
          crx[lv] = cr;
          xfx[lv] = cr * dxa[m] * gdyp[tile * gdy + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)] *
                    sg[tile * 9 * gs2 + (w > T(0) ? 2 : 0) * gs2 + (long long)(j - d.jsd) * d.nid + (iu - d.isd)];
          cx[lv] = first ? T(0) + cr : cx[lv] + cr;
        }, d.is, d.ie + 1, d.jsd, d.jed, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const long long lv = (long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
          const T w = v[lv];
          const int ju = w > T(0) ? j - 1 : j;
          const long long m = tile * gs2 + (long long)(ju - d.jsd) * d.nid + (i - d.isd);
          const T cr = bdt * w / dya[m];
          cry[lv] = cr;
          yfx[lv] = cr * dya[m] * gdxp[tile * gdx + (long long)(j - d.jsd) * d.nid + (i - d.isd)] *
                    sg[tile * 9 * gs2 + (w > T(0) ? 3 : 1) * gs2 + (long long)(ju - d.jsd) * d.nid + (i - d.isd)];
          cy[lv] = first ? T(0) + cr : cy[lv] + cr;
        }, d.isd, d.ied, d.js, d.je + 1, nz))
      return 1;
    // ra_x, ra_y
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *xk = xfx + (long long)z * d.sx + (long long)(j - d.jsd) * d.nix;
          rax[(long long)z * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
              area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + xk[i - d.is] - xk[i + 1 - d.is];
        }, d.is, d.ie, d.jsd, d.jed, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *yk = yfx + (long long)z * d.sy;
          ray[(long long)z * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
              area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] +
              yk[(long long)(j - d.js) * d.nid + (i - d.isd)] - yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
        }, d.isd, d.ied, d.js, d.je, nz))
      return 1;
    // fv_tp_2d of delp
    if (cc(2)) return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          ColQ<T> qa{delp + (long long)z * d.s2 + (i - d.isd), d.jsd, d.nid};
          ColQ<T> da{dya + tile * gs2 + (i - d.isd), d.jsd, d.nid};
          const long long lv = (long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
          t_fy2[lv] = ppm_flux<T>(qa, da, j, cry[lv], gy, ord_in);
        }, d.isd, d.ied, d.js, d.je + 1, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *yk = yfx + (long long)z * d.sy, *f2 = t_fy2 + (long long)z * d.sy;
          const T fa = yk[(long long)(j - d.js) * d.nid + (i - d.isd)] * f2[(long long)(j - d.js) * d.nid + (i - d.isd)];
          const T fb = yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)] * f2[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
          t_qi[(long long)z * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
              (delp[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] *
                   area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
              ray[(long long)z * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)];
        }, d.isd, d.ied, d.js, d.je, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          RowQ<T> qa{t_qi + (long long)z * d.sry + (long long)(j - d.js) * d.nid, d.isd};
          RowQ<T> da{dxa + tile * gs2 + (long long)(j - d.jsd) * d.nid, d.isd};
          const T cv = crx[(long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)];
          fx[(long long)z * d.smx + (long long)(j - d.js) * d.nix + (i - d.is)] = ppm_flux<T>(qa, da, i, cv, gx, ord_ou);
        }, d.is, d.ie + 1, d.js, d.je, nz))
      return 1;
    if (cc(1)) return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          RowQ<T> qa{delp + (long long)z * d.s2 + (long long)(j - d.jsd) * d.nid, d.isd};
          RowQ<T> da{dxa + tile * gs2 + (long long)(j - d.jsd) * d.nid, d.isd};
          const long long lv = (long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
          t_fx2[lv] = ppm_flux<T>(qa, da, i, crx[lv], gx, ord_in);
        }, d.is, d.ie + 1, d.jsd, d.jed, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *xk = xfx + (long long)z * d.sx + (long long)(j - d.jsd) * d.nix;
          const T *f2 = t_fx2 + (long long)z * d.sx + (long long)(j - d.jsd) * d.nix;
          const T fa = xk[i - d.is] * f2[i - d.is];
          const T fb = xk[i + 1 - d.is] * f2[i + 1 - d.is];
          t_qj[(long long)z * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
              (delp[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] *
                   area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
              rax[(long long)z * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)];
        }, d.is, d.ie, d.jsd, d.jed, nz))
      return 1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          ColQ<T> qa{t_qj + (long long)z * d.srx + (i - d.is), d.jsd, d.niy};
          ColQ<T> da{dya + tile * gs2 + (i - d.isd), d.jsd, d.nid};
          const T cv = cry[(long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)];
          fy[(long long)z * d.smy + (long long)(j - d.js) * d.niy + (i - d.is)] = ppm_flux<T>(qa, da, j, cv, gy, ord_ou);
        }, d.is, d.ie, d.js, d.je + 1, nz))
      return 1;
    // flux averaging with the area fluxes; mfx, mfy accumulate
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          if (i <= d.ie + 1 && j <= d.je) {
            const long long o = (long long)z * d.smx + (long long)(j - d.js) * d.nix + (i - d.is);
            const long long ox = (long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
            const T f = T(0.5) * (fx[o] + t_fx2[ox]) * xfx[ox];
            fx[o] = f;
            mfx[o] = first ? T(0) + f : mfx[o] + f;
          }
          if (i <= d.ie && j <= d.je + 1) {
            const long long o = (long long)z * d.smy + (long long)(j - d.js) * d.niy + (i - d.is);
            const long long oy = (long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
            const T f = T(0.5) * (fy[o] + t_fy2[oy]) * yfx[oy];
            fy[o] = f;
            mfy[o] = first ? T(0) + f : mfy[o] + f;
          }
        }, d.is, d.ie + 1, d.js, d.je + 1, nz))
      return 1;
    // delp by the flux divergence
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *fxr = fx + (long long)z * d.smx + (long long)(j - d.js) * d.nix;
          const T *fyc = fy + (long long)z * d.smy;
          T &p = delp[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
          p = p + ((fxr[i - d.is] - fxr[i + 1 - d.is]) +
                   (fyc[(long long)(j - d.js) * d.niy + (i - d.is)] - fyc[(long long)(j + 1 - d.js) * d.niy + (i - d.is)])) *
                      rarea[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
        }, d.is, d.ie, d.js, d.je, nz))
      return 1;
    if (nt > 1 && it + 1 < n_split)
      if (halo_fill<T>(c, delp, (long long)npz * d.s2, nullptr, 0, npz)) return 1;
  }
  return 0;
}

}  // namespace fv3

using namespace fv3;

extern "C" int fv3_dyn_core_fluxes(fv3_ctx *h, int form, int n_split, double bdt, int hord_dp, const void *uc,
                                   const void *vc, void *delp, void *dp1, void *cx, void *cy, void *mfx, void *mfy,
                                   void *stream) {
  if (!h || !uc || !vc || !delp || !dp1 || !cx || !cy || !mfx || !mfy)
    return set_error("fv3_dyn_core_fluxes: NULL argument");
  if (n_split < 1) return set_error("n_split must be >= 1 (got %d)", n_split);
  if (hord_dp < 7 || hord_dp > 13) return set_error("hord_dp=%d: only the monotone orders 7..13 are implemented", hord_dp);
  if (!(bdt > 0.0)) return set_error("bdt must be > 0 (got %g)", bdt);
  if (delp == dp1) return set_error("fv3_dyn_core_fluxes: delp and dp1 must be distinct arrays");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.grid_set) return set_error("grid metrics not set (fv3_ctx_set_grid)");
  hipStream_t s = stream ? (hipStream_t)stream : c.stream;
  int rc;
  if (form == FV3_FORM_PLAIN64)
    rc = dyn_core_fluxes<double>(c, n_split, bdt, hord_dp, (const double *)uc, (const double *)vc, (double *)delp,
                                 (double *)dp1, (double *)cx, (double *)cy, (double *)mfx, (double *)mfy, s);
  else if (form == FV3_FORM_FUSED32)
    rc = dyn_core_fluxes<float>(c, n_split, (float)bdt, hord_dp, (const float *)uc, (const float *)vc,
                                (float *)delp, (float *)dp1, (float *)cx, (float *)cy, (float *)mfx, (float *)mfy, s);
  else
    rc = set_error("form %d: the dyn_core fluxes run FV3_FORM_PLAIN64 (double) or FV3_FORM_FUSED32 (float)", form);
  if (rc) return 1;
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}
