// tlad.hip -- tangent-linear (TL) and adjoint (AD) of the advection + tracer
// remap step in FP64: tracer_2d (one nsplit, q_split forcing) or
// tracer_2d_1L (per-level nsplit) over fv_tp_2d, then the tracer part of
// Lagrangian_to_Eulerian (cs_profile/scalar_profile iv=0 and the layer
// integrals of mapn_tracer / map1_q2).  Upstream: fv_tracer2d.F90,
// tp_core.F90, fv_mapz.F90 of GFDL_atmos_cubed_sphere (empty submodule,
// .gitmodules:1-4).
//
// The derivative is taken with respect to the interior tracer values, with
// every limiter branch, upwind choice and source-layer search frozen at the
// forward trajectory; winds, thicknesses and the (frozen, single-tile) halo
// are fixed.  TL evaluates the forward stages on Dual<1> numbers: the value
// part reproduces the plain FP64 forward bit for bit, the tangent part is the
// frozen-branch derivative.  AD recomputes each sub-cycle from its
// checkpointed input state, takes the local Jacobian of every PPM face flux
// (6 stencil values, Dual<6>) and of every remapped cell's limited parabola
// (7 local values, Dual<7>), and applies the transposes of the stages in
// reverse order: flux-form update, flux averaging, the four sweeps and the
// two advective updates of fv_tp_2d; the layer integrals, the limiters and
// the interface tridiagonal solve of the remap.
#include <algorithm>
#include <vector>

#include "dual.h"
#include "fused.h"
#include "fv3_internal.h"
#include "mapz_column.h"

namespace fv3 {

namespace {

using D6 = Dual<6>;
using D7 = Dual<7>;

template <typename F>
__global__ void __launch_bounds__(256) k_box(F f, int i0, int i1, int j0, int j1) {
  const int ni = i1 - i0 + 1;
  const long long n = (long long)ni * (j1 - j0 + 1);
  const int z = blockIdx.z;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    f(i0 + (int)(t % ni), j0 + (int)(t / ni), z);
}

template <typename F>
int box(hipStream_t s, F f, int i0, int i1, int j0, int j1, int nz) {
  if (i1 < i0 || j1 < j0 || nz <= 0) return 0;
  const long long n = (long long)(i1 - i0 + 1) * (j1 - j0 + 1);
  dim3 g((unsigned)std::min<long long>((n + 255) / 256, 2048), 1, (unsigned)nz);
  hipLaunchKernelGGL(k_box<F>, g, dim3(256), 0, s, f, i0, i1, j0, j1);
  FV3_HIP(hipGetLastError());
  return 0;
}

// ---- scalar load/store over (value, tangent) array pairs
template <typename S>
struct Fld;
template <>
struct Fld<double> {
  __device__ static double ld(const double *v, const double *, long long o) { return v[o]; }
  __device__ static void st(double *v, double *, long long o, double x) { v[o] = x; }
};
template <>
struct Fld<D1> {
  __device__ static D1 ld(const double *v, const double *t, long long o) {
    D1 r(v[o]);
    r.d[0] = t[o];
    return r;
  }
  __device__ static void st(double *v, double *t, long long o, const D1 &x) {
    v[o] = x.v;
    t[o] = x.d[0];
  }
};

// (isd:ied, jsd:jed) working field along a row / column, with the
// copy_corners substitution of the given direction in the corner blocks
template <typename S>
struct FieldLine {
  const double *v, *t;
  int fixed;  // j of a row, i of a column
  bool row;
  int isd, jsd, nid, npx, npy, dir;
  unsigned corners;
  __device__ S operator()(int m) const {
    int i = row ? m : fixed, j = row ? fixed : m;
    int si, sj;
    if (corners && corner_src(i, j, npx, npy, dir, corners, si, sj)) {
      i = si;
      j = sj;
    }
    return Fld<S>::ld(v, t, (long long)(j - jsd) * nid + (i - isd));
  }
};

// line of an S-typed array: element m at p[(m - lo) * stride]
template <typename S>
struct Line {
  const S *p;
  int lo;
  long long stride;
  __device__ S operator()(int m) const { return p[(long long)(m - lo) * stride]; }
};

// metric line (double)
struct MLine {
  const double *p;
  int lo;
  long long stride;
  __device__ double operator()(int m) const { return p[(long long)(m - lo) * stride]; }
};

// seeds the 6 stencil values of face f (cells f-3 .. f+2) of a double line
template <typename Q>
struct Seed6 {
  Q q;
  int f;
  __device__ D6 operator()(int m) const {
    D6 r(q(m));
    const int o = m - (f - 3);
    if (o >= 0 && o < 6) r.d[o] = 1.0;
    return r;
  }
};

// per-level wind / area-flux data shared by every tracer of a level
struct LevelData {
  const double *cx, *cy, *xfx, *yfx, *rax, *ray, *mfx, *mfy;  // scaled, per level
};

struct Geo {
  Dims d;
  const double *area, *rarea, *dxa, *dya;
  int npz;
};

// Temporaries of one slab (S-typed), strides from Dims.
template <typename S>
struct SubTemps {
  S *fy2, *qi, *fx, *fx2, *qj, *fy;
};

// ---- forward stages of one sub-cycle over the active slabs of a chunk.
// slab p -> field slab z = zs[p] (iq*npz + k); work field (wv, wt) updated
// in place on the interior.
template <typename S>
int subcycle_forward(hipStream_t s, const Geo &G, const LevelData &L, const int *zs, int np,
                     double *wv, double *wt, const SubTemps<S> &T, const double *dpa,
                     const double *dpb, int ord_in, int ord_ou, bool update) {
  const Dims d = G.d;
  const int npz = G.npz;
  const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
  const unsigned corners = d.bounded ? 0u : d.corners;
  const double *area = G.area, *rarea = G.rarea, *dxa = G.dxa, *dya = G.dya;
  // fy2 = yppm(q with dir-2 corners, ord_in) over (isd:ied, js:je+1)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        FieldLine<S> qa{wv + (long long)z * d.s2, wt ? wt + (long long)z * d.s2 : nullptr, i, false,
                        d.isd, d.jsd, d.nid, d.npx, d.npy, 2, corners};
        MLine da{dya + (i - d.isd), d.jsd, d.nid};
        const S cv = S(L.cy[(long long)k * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]);
        T.fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)] = ppm_flux<S>(qa, da, j, cv, gy, ord_in);
      }, d.isd, d.ied, d.js, d.je + 1, np))
    return -1;
  // q_i over (isd:ied, js:je)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        const double *yk = L.yfx + (long long)k * d.sy;
        const S *f2 = T.fy2 + (long long)p * d.sy;
        const S fa = S(yk[(long long)(j - d.js) * d.nid + (i - d.isd)]) * f2[(long long)(j - d.js) * d.nid + (i - d.isd)];
        const S fb = S(yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)]) * f2[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
        const long long o = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        const S w = Fld<S>::ld(wv + (long long)z * d.s2, wt ? wt + (long long)z * d.s2 : nullptr, o);
        T.qi[(long long)p * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
            (w * S(area[o]) + fa - fb) / S(L.ray[(long long)k * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)]);
      }, d.isd, d.ied, d.js, d.je, np))
    return -1;
  // fx = xppm(q_i, ord_ou) over (is:ie+1, js:je)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        Line<S> qa{T.qi + (long long)p * d.sry + (long long)(j - d.js) * d.nid, d.isd, 1};
        MLine da{dxa + (long long)(j - d.jsd) * d.nid, d.isd, 1};
        const S cv = S(L.cx[(long long)k * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]);
        T.fx[(long long)p * d.smx + (long long)(j - d.js) * d.nix + (i - d.is)] = ppm_flux<S>(qa, da, i, cv, gx, ord_ou);
      }, d.is, d.ie + 1, d.js, d.je, np))
    return -1;
  // fx2 = xppm(q with dir-1 corners, ord_in) over (is:ie+1, jsd:jed)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        FieldLine<S> qa{wv + (long long)z * d.s2, wt ? wt + (long long)z * d.s2 : nullptr, j, true,
                        d.isd, d.jsd, d.nid, d.npx, d.npy, 1, corners};
        MLine da{dxa + (long long)(j - d.jsd) * d.nid, d.isd, 1};
        const S cv = S(L.cx[(long long)k * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]);
        T.fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)] = ppm_flux<S>(qa, da, i, cv, gx, ord_in);
      }, d.is, d.ie + 1, d.jsd, d.jed, np))
    return -1;
  // q_j over (is:ie, jsd:jed)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        const double *xk = L.xfx + (long long)k * d.sx + (long long)(j - d.jsd) * d.nix;
        const S *f2 = T.fx2 + (long long)p * d.sx + (long long)(j - d.jsd) * d.nix;
        const S fa = S(xk[i - d.is]) * f2[i - d.is];
        const S fb = S(xk[i + 1 - d.is]) * f2[i + 1 - d.is];
        const long long o = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        const S w = Fld<S>::ld(wv + (long long)z * d.s2, wt ? wt + (long long)z * d.s2 : nullptr, o);
        T.qj[(long long)p * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
            (w * S(area[o]) + fa - fb) / S(L.rax[(long long)k * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)]);
      }, d.is, d.ie, d.jsd, d.jed, np))
    return -1;
  // fy = yppm(q_j, ord_ou) over (is:ie, js:je+1)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        Line<S> qa{T.qj + (long long)p * d.srx + (i - d.is), d.jsd, d.niy};
        MLine da{dya + (i - d.isd), d.jsd, d.nid};
        const S cv = S(L.cy[(long long)k * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]);
        T.fy[(long long)p * d.smy + (long long)(j - d.js) * d.niy + (i - d.is)] = ppm_flux<S>(qa, da, j, cv, gy, ord_ou);
      }, d.is, d.ie, d.js, d.je + 1, np))
    return -1;
  if (!update) return 0;
  // flux averaging with the mass fluxes
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        if (i <= d.ie + 1 && j <= d.je) {
          const long long o = (long long)(j - d.js) * d.nix + (i - d.is);
          S &f = T.fx[(long long)p * d.smx + o];
          f = S(0.5) * (f + T.fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]) * S(L.mfx[(long long)k * d.smx + o]);
        }
        if (i <= d.ie && j <= d.je + 1) {
          const long long o = (long long)(j - d.js) * d.niy + (i - d.is);
          S &f = T.fy[(long long)p * d.smy + o];
          f = S(0.5) * (f + T.fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]) * S(L.mfy[(long long)k * d.smy + o]);
        }
      }, d.is, d.ie + 1, d.js, d.je + 1, np))
    return -1;
  // flux-form update of the interior, in place (each cell reads only itself)
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        const long long o = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        const S *fxr = T.fx + (long long)p * d.smx + (long long)(j - d.js) * d.nix;
        const S *fyc = T.fy + (long long)p * d.smy;
        double *v = wv + (long long)z * d.s2;
        double *t = wt ? wt + (long long)z * d.s2 : nullptr;
        const S w = Fld<S>::ld(v, t, o);
        const S val = (w * S(dpa[(long long)k * d.s2 + o]) +
                       (fxr[i - d.is] - fxr[i + 1 - d.is] + fyc[(long long)(j - d.js) * d.niy + (i - d.is)] -
                        fyc[(long long)(j + 1 - d.js) * d.niy + (i - d.is)]) *
                           S(rarea[o])) /
                      S(dpb[(long long)k * d.s2 + o]);
        Fld<S>::st(v, t, o, val);
      }, d.is, d.ie, d.js, d.je, np))
    return -1;
  return 0;
}

// Local Jacobians (6 weights per face) of the four sweeps of one sub-cycle,
// from the primal temporaries.
struct SweepW {
  double *fy2, *fx, *fx2, *fy;  // 6 per face, face layouts of sy, smx, sx, smy
};

int sweep_weights(hipStream_t s, const Geo &G, const LevelData &L, const int *zs, int np,
                  const double *wv, const SubTemps<double> &T, const SweepW &W, int ord_in,
                  int ord_ou) {
  const Dims d = G.d;
  const int npz = G.npz;
  const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
  const unsigned corners = d.bounded ? 0u : d.corners;
  const double *dxa = G.dxa, *dya = G.dya;
  auto put = [] __device__(double *w, long long face, const D6 &f) {
    for (int n = 0; n < 6; n++) w[face * 6 + n] = f.d[n];
  };
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        FieldLine<double> base{wv + (long long)z * d.s2, nullptr, i, false, d.isd, d.jsd, d.nid, d.npx, d.npy, 2, corners};
        Seed6<FieldLine<double>> qa{base, j};
        MLine da{dya + (i - d.isd), d.jsd, d.nid};
        const D6 cv(L.cy[(long long)k * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]);
        put(W.fy2, (long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd), ppm_flux<D6>(qa, da, j, cv, gy, ord_in));
      }, d.isd, d.ied, d.js, d.je + 1, np))
    return -1;
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        Line<double> base{T.qi + (long long)p * d.sry + (long long)(j - d.js) * d.nid, d.isd, 1};
        Seed6<Line<double>> qa{base, i};
        MLine da{dxa + (long long)(j - d.jsd) * d.nid, d.isd, 1};
        const D6 cv(L.cx[(long long)k * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]);
        put(W.fx, (long long)p * d.smx + (long long)(j - d.js) * d.nix + (i - d.is), ppm_flux<D6>(qa, da, i, cv, gx, ord_ou));
      }, d.is, d.ie + 1, d.js, d.je, np))
    return -1;
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        FieldLine<double> base{wv + (long long)z * d.s2, nullptr, j, true, d.isd, d.jsd, d.nid, d.npx, d.npy, 1, corners};
        Seed6<FieldLine<double>> qa{base, i};
        MLine da{dxa + (long long)(j - d.jsd) * d.nid, d.isd, 1};
        const D6 cv(L.cx[(long long)k * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]);
        put(W.fx2, (long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is), ppm_flux<D6>(qa, da, i, cv, gx, ord_in));
      }, d.is, d.ie + 1, d.jsd, d.jed, np))
    return -1;
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        Line<double> base{T.qj + (long long)p * d.srx + (i - d.is), d.jsd, d.niy};
        Seed6<Line<double>> qa{base, j};
        MLine da{dya + (i - d.isd), d.jsd, d.nid};
        const D6 cv(L.cy[(long long)k * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]);
        put(W.fy, (long long)p * d.smy + (long long)(j - d.js) * d.niy + (i - d.is), ppm_flux<D6>(qa, da, j, cv, gy, ord_ou));
      }, d.is, d.ie, d.js, d.je + 1, np))
    return -1;
  return 0;
}

// Adjoint temporaries of one slab
struct AdjTemps {
  double *qa;              // (isd:ied, jsd:jed) adjoint of the sub-cycle input
  double *Fx, *Fy;         // averaged-flux adjoints (smx, smy)
  double *fx, *fx2, *fy, *fy2;  // sweep-output adjoints (smx, sx, smy, sy)
  double *qj, *qi;         // (srx, sry)
  double *cg;              // 36 per slab: adjoints read through the corner blocks
};

// corner-block slot (0..35) of cell (i, j): block 0 sw, 1 se, 2 ne, 3 nw, then
// 3 x 3 cells i fastest; -1 outside the corner blocks
__device__ __forceinline__ int corner_slot(int i, int j, int npx, int npy) {
  const bool ilo = i <= 0 && i >= 1 - NG, ihi = i >= npx && i <= npx + NG - 1;
  const bool jlo = j <= 0 && j >= 1 - NG, jhi = j >= npy && j <= npy + NG - 1;
  int blk;
  if (ilo && jlo) blk = 0;
  else if (ihi && jlo) blk = 1;
  else if (ihi && jhi) blk = 2;
  else if (ilo && jhi) blk = 3;
  else return -1;
  const int a = ilo ? i - (1 - NG) : i - npx, b = jlo ? j - (1 - NG) : j - npy;
  return blk * 9 + b * 3 + a;
}

// Reverse of one sub-cycle on the active slabs of a chunk: qbar (interior of
// each slab) is replaced by the adjoint of the sub-cycle input.
// With full (six faces) the adjoint of every cell the sub-cycle read is kept
// in qbar, halo cells included (the halo transposes it onto the neighbour
// faces afterwards); an adjoint read through a corner-block cell goes to the
// cell copy_corners took its value from.  Without it (one tile, frozen
// halos) only the interior is written.
int subcycle_adjoint(hipStream_t s, const Geo &G, const LevelData &L, const int *zs, int np,
                     double *qbar, const SweepW &W, const AdjTemps &A, const double *dpa,
                     const double *dpb, bool full) {
  const Dims d = G.d;
  const int npz = G.npz;
  const double *area = G.area, *rarea = G.rarea;
  const unsigned corners = d.bounded ? 0u : d.corners;
  // corner-block adjoints of sweep direction dir go to their sources
  auto redirect = [&](int dir) {
    return box(s, [=] __device__(int slot, int, int p) {
          const int blk = slot / 9, cc = slot % 9;
          const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + cc % 3;
          const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + cc / 3;
          int si, sj;
          if (corner_src(i, j, d.npx, d.npy, dir, corners, si, sj))
            A.qa[(long long)p * d.s2 + (long long)(sj - d.jsd) * d.nid + (si - d.isd)] += A.cg[(long long)p * 36 + slot];
        }, 0, 35, 0, 0, np);
  };
  // update: qa = qbar dp1/dp2 on the interior, 0 elsewhere
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        const long long o = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        const bool in = i >= d.is && i <= d.ie && j >= d.js && j <= d.je;
        A.qa[(long long)p * d.s2 + o] =
            in ? qbar[(long long)z * d.s2 + o] * dpa[(long long)k * d.s2 + o] / dpb[(long long)k * d.s2 + o] : 0.0;
      }, d.isd, d.ied, d.jsd, d.jed, np))
    return -1;
  // averaged-flux adjoints: r = qbar rarea / dp2 on the interior
  if (box(s, [=] __device__(int i, int j, int p) {
        const int z = zs[p], k = z % npz;
        auto r = [&](int ii, int jj) -> double {
          if (ii < d.is || ii > d.ie || jj < d.js || jj > d.je) return 0.0;
          const long long o = (long long)(jj - d.jsd) * d.nid + (ii - d.isd);
          return qbar[(long long)z * d.s2 + o] * rarea[o] / dpb[(long long)k * d.s2 + o];
        };
        if (i <= d.ie + 1 && j <= d.je) {
          const long long o = (long long)(j - d.js) * d.nix + (i - d.is);
          const double Fb = r(i, j) - r(i - 1, j);
          A.Fx[(long long)p * d.smx + o] = Fb;
          const double h = 0.5 * Fb * L.mfx[(long long)k * d.smx + o];
          A.fx[(long long)p * d.smx + o] = h;
          A.fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)] = h;
        }
        if (i <= d.ie && j <= d.je + 1) {
          const long long o = (long long)(j - d.js) * d.niy + (i - d.is);
          const double Fb = r(i, j) - r(i, j - 1);
          A.Fy[(long long)p * d.smy + o] = Fb;
          const double h = 0.5 * Fb * L.mfy[(long long)k * d.smy + o];
          A.fy[(long long)p * d.smy + o] = h;
          A.fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)] = h;
        }
      }, d.is, d.ie + 1, d.js, d.je + 1, np))
    return -1;
  // fx2 adjoint outside rows js..je and fy2 adjoint outside columns is..ie start at 0
  if (box(s, [=] __device__(int i, int j, int p) {
        if (j < d.js || j > d.je) A.fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)] = 0.0;
      }, d.is, d.ie + 1, d.jsd, d.jed, np))
    return -1;
  if (box(s, [=] __device__(int i, int j, int p) {
        if (i < d.is || i > d.ie) A.fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)] = 0.0;
      }, d.isd, d.ied, d.js, d.je + 1, np))
    return -1;
  // outer y-sweep: qj_bar(i, m) = sum over faces f = m-2..m+3 of W_fy(i, f)[m-f+3] fy_bar(i, f)
  if (box(s, [=] __device__(int i, int m, int p) {
        double acc = 0.0;
        for (int f = m - 2; f <= m + 3; f++) {
          if (f < d.js || f > d.je + 1) continue;
          const long long face = (long long)p * d.smy + (long long)(f - d.js) * d.niy + (i - d.is);
          acc += W.fy[face * 6 + (m - f + 3)] * A.fy[face];
        }
        A.qj[(long long)p * d.srx + (long long)(m - d.jsd) * d.niy + (i - d.is)] = acc;
      }, d.is, d.ie, d.jsd, d.jed, np))
    return -1;
  // q_j: b = qj_bar / ra_x; qa += b area; fx2_bar(f) += xfx(f) (b(f) - b(f-1))
  if (box(s, [=] __device__(int i, int j, int p) {
        const int k = zs[p] % npz;
        if (i <= d.ie) {
          const long long o = (long long)(j - d.jsd) * d.niy + (i - d.is);
          const double b = A.qj[(long long)p * d.srx + o] / L.rax[(long long)k * d.srx + o];
          const long long oq = (long long)(j - d.jsd) * d.nid + (i - d.isd);
          A.qa[(long long)p * d.s2 + oq] += b * area[oq];
        }
      }, d.is, d.ie, d.jsd, d.jed, np))
    return -1;
  if (box(s, [=] __device__(int f, int j, int p) {
        const int k = zs[p] % npz;
        auto b = [&](int ii) -> double {
          if (ii < d.is || ii > d.ie) return 0.0;
          const long long o = (long long)(j - d.jsd) * d.niy + (ii - d.is);
          return A.qj[(long long)p * d.srx + o] / L.rax[(long long)k * d.srx + o];
        };
        const long long o = (long long)(j - d.jsd) * d.nix + (f - d.is);
        A.fx2[(long long)p * d.sx + o] += L.xfx[(long long)k * d.sx + o] * (b(f) - b(f - 1));
      }, d.is, d.ie + 1, d.jsd, d.jed, np))
    return -1;
  // inner x-sweep: qa(m, j) += sum of W_fx2(f, j)[m-f+3] fx2_bar(f, j) (the
  // interior; with full every cell, corner-block reads set aside in cg)
  if (box(s, [=] __device__(int m, int j, int p) {
        double acc = 0.0;
        for (int f = m - 2; f <= m + 3; f++) {
          if (f < d.is || f > d.ie + 1) continue;
          const long long face = (long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (f - d.is);
          acc += W.fx2[face * 6 + (m - f + 3)] * A.fx2[face];
        }
        int si, sj;
        if (full && corners && corner_src(m, j, d.npx, d.npy, 1, corners, si, sj)) {
          A.cg[(long long)p * 36 + corner_slot(m, j, d.npx, d.npy)] = acc;
          return;
        }
        A.qa[(long long)p * d.s2 + (long long)(j - d.jsd) * d.nid + (m - d.isd)] += acc;
      }, full ? d.isd : d.is, full ? d.ied : d.ie, full ? d.jsd : d.js, full ? d.jed : d.je, np))
    return -1;
  if (full && corners && redirect(1)) return -1;
  // outer x-sweep: qi_bar(m, j) = sum of W_fx(f, j)[m-f+3] fx_bar(f, j)
  if (box(s, [=] __device__(int m, int j, int p) {
        double acc = 0.0;
        for (int f = m - 2; f <= m + 3; f++) {
          if (f < d.is || f > d.ie + 1) continue;
          const long long face = (long long)p * d.smx + (long long)(j - d.js) * d.nix + (f - d.is);
          acc += W.fx[face * 6 + (m - f + 3)] * A.fx[face];
        }
        A.qi[(long long)p * d.sry + (long long)(j - d.js) * d.nid + (m - d.isd)] = acc;
      }, d.isd, d.ied, d.js, d.je, np))
    return -1;
  // q_i: c = qi_bar / ra_y; qa += c area; fy2_bar(i, f) += yfx(f) (c(f) - c(f-1))
  if (box(s, [=] __device__(int i, int j, int p) {
        const int k = zs[p] % npz;
        const long long o = (long long)(j - d.js) * d.nid + (i - d.isd);
        const double c = A.qi[(long long)p * d.sry + o] / L.ray[(long long)k * d.sry + o];
        const long long oq = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        A.qa[(long long)p * d.s2 + oq] += c * area[oq];
      }, d.isd, d.ied, d.js, d.je, np))
    return -1;
  if (box(s, [=] __device__(int i, int f, int p) {
        const int k = zs[p] % npz;
        auto c = [&](int jj) -> double {
          if (jj < d.js || jj > d.je) return 0.0;
          const long long o = (long long)(jj - d.js) * d.nid + (i - d.isd);
          return A.qi[(long long)p * d.sry + o] / L.ray[(long long)k * d.sry + o];
        };
        const long long o = (long long)(f - d.js) * d.nid + (i - d.isd);
        A.fy2[(long long)p * d.sy + o] += L.yfx[(long long)k * d.sy + o] * (c(f) - c(f - 1));
      }, d.isd, d.ied, d.js, d.je + 1, np))
    return -1;
  // inner y-sweep: qa(i, m) += sum of W_fy2(i, f)[m-f+3] fy2_bar(i, f); then
  // qbar <- qa on the interior (with full: on every cell)
  if (box(s, [=] __device__(int i, int m, int p) {
        double acc = 0.0;
        for (int f = m - 2; f <= m + 3; f++) {
          if (f < d.js || f > d.je + 1) continue;
          const long long face = (long long)p * d.sy + (long long)(f - d.js) * d.nid + (i - d.isd);
          acc += W.fy2[face * 6 + (m - f + 3)] * A.fy2[face];
        }
        const long long o = (long long)(m - d.jsd) * d.nid + (i - d.isd);
        if (!full) {
          qbar[(long long)zs[p] * d.s2 + o] = A.qa[(long long)p * d.s2 + o] + acc;
          return;
        }
        int si, sj;
        if (corners && corner_src(i, m, d.npx, d.npy, 2, corners, si, sj)) {
          A.cg[(long long)p * 36 + corner_slot(i, m, d.npx, d.npy)] = acc;
          return;
        }
        A.qa[(long long)p * d.s2 + o] += acc;
      }, full ? d.isd : d.is, full ? d.ied : d.ie, full ? d.jsd : d.js, full ? d.jed : d.je, np))
    return -1;
  if (!full) return 0;
  if (corners && redirect(2)) return -1;
  if (box(s, [=] __device__(int i, int j, int p) {
        const long long o = (long long)(j - d.jsd) * d.nid + (i - d.isd);
        qbar[(long long)zs[p] * d.s2 + o] = A.qa[(long long)p * d.s2 + o];
      }, d.isd, d.ied, d.jsd, d.jed, np))
    return -1;
  return 0;
}

// Transposed edge-halo refresh of six stacked tiles: every interior cell adds
// the adjoints of the halo cells that copied it (in the table's order, so the
// result does not depend on scheduling), then every cell outside the
// interior is cleared.  usrc: the distinct source cells; ptr/dst: the halo
// cells of each (tile << 40 | cell, as the forward table).
__global__ void __launch_bounds__(256) k_halo_adjoint(double *q, long long tile_stride, long long s2,
                                                      const long long *usrc, const int *ptr, const long long *dst,
                                                      int nsrc) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nsrc) return;
  const long long sl = blockIdx.y;
  double acc = 0.0;
  for (int k = ptr[e]; k < ptr[e + 1]; k++) {
    const long long td = dst[k] >> 40, cd = dst[k] & ((1ll << 40) - 1);
    acc += q[td * tile_stride + sl * s2 + cd];
  }
  const long long ts = usrc[e] >> 40, cs = usrc[e] & ((1ll << 40) - 1);
  double &v = q[ts * tile_stride + sl * s2 + cs];
  v = v + acc;
}

// ---------------------------------------------------------------- remap
struct RemapTL {
  int km, nq, ni, is, js, isd, jsd, nid, kord;
  long long ncol, s2;
  const double *ak, *bk;
  double ptop;
  const double *dp1;
  double *delp_out;
  const double *qv, *qt;  // transported state and tangent
  double *ov, *ot;        // remapped state and tangent
  // scratch [level][column]
  double *pe1, *pe2, *d1, *d2;
  D1 *a4, *qi, *gam, *dm;
  uint8_t *flags;
};

// map_col with the partial sums carried in S (values follow the upstream
// expressions of map1_q2 / mapn_tracer)
template <typename S, typename OUT>
__device__ void map_col_s(const CV<double> &PE1, const CV<double> &PE2, const CV<double> &D1v,
                          const CV<S> &A1, const CV<S> &A2, const CV<S> &A3, const CV<S> &A4, int km,
                          bool mapn, const CV<double> &DEN, OUT out) {
  const double r3 = 1.0 / 3.0, r23 = 2.0 / 3.0;
  int k0 = 1;
  for (int k = 1; k <= km; k++) {
    S qsum(0.0);
    bool single = false;
    for (int l = k0; l <= km; l++) {
      if (PE2(k) >= PE1(l) && PE2(k) <= PE1(l + 1)) {
        const double pl = (PE2(k) - PE1(l)) / D1v(l);
        if (PE2(k + 1) <= PE1(l + 1)) {
          const double pr = (PE2(k + 1) - PE1(l)) / D1v(l);
          S v;
          if (mapn) {
            double fac1 = pr + pl;
            double fac2 = r3 * (pr * fac1 + pl * pl);
            fac1 = 0.5 * fac1;
            v = A2(l) + (A4(l) + A3(l) - A2(l)) * S(fac1) - A4(l) * S(fac2);
          } else {
            v = A2(l) + S(0.5) * (A4(l) + A3(l) - A2(l)) * S(pr + pl) - A4(l) * S(r3) * S(pr * (pr + pl) + pl * pl);
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
          qsum = S(dp) * (A2(l) + (A4(l) + A3(l) - A2(l)) * S(fac1) - A4(l) * S(fac2));
        } else {
          qsum = S(PE1(l + 1) - PE2(k)) *
                 (A2(l) + S(0.5) * (A4(l) + A3(l) - A2(l)) * S(1.0 + pl) - A4(l) * S(r3 * (1.0 + pl * (1.0 + pl))));
        }
        for (int m = l + 1; m <= km; m++) {
          if (PE2(k + 1) > PE1(m + 1)) {
            qsum = qsum + S(D1v(m)) * A1(m);
          } else {
            const double dp = PE2(k + 1) - PE1(m);
            const double esl = dp / D1v(m);
            if (mapn) {
              const double f1 = 0.5 * esl, f2 = 1.0 - r23 * esl;
              qsum = qsum + S(dp) * (A2(m) + S(f1) * (A3(m) - A2(m) + A4(m) * S(f2)));
            } else {
              qsum = qsum + S(dp) * (A2(m) + S(0.5 * esl) * (A3(m) - A2(m) + A4(m) * S(1.0 - r23 * esl)));
            }
            k0 = m;
            break;
          }
        }
        break;
      }
    }
    if (!single) out(k, qsum / S(DEN(k)));
  }
}

// pressures and thicknesses of one column (as the Lagrangian_to_Eulerian subset)
__device__ void column_pressures(int km, double ptop, const double *ak, const double *bk,
                                 const double *dl, long long s2, const CV<double> &PE1,
                                 const CV<double> &PE2, const CV<double> &D1v, const CV<double> &D2v) {
  PE1(1) = ptop;
  for (int k = 1; k <= km; k++) PE1(k + 1) = PE1(k) + dl[(long long)(k - 1) * s2];
  PE2(1) = ptop;
  PE2(km + 1) = PE1(km + 1);
  for (int k = 2; k <= km; k++) PE2(k) = ak[k - 1] + bk[k - 1] * PE1(km + 1);
  for (int k = 1; k <= km; k++) {
    D2v(k) = PE2(k + 1) - PE2(k);
    D1v(k) = PE1(k + 1) - PE1(k);
  }
}

__global__ void __launch_bounds__(64) k_remap_tl(const RemapTL a) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  const long long nc = a.ncol;
  const int km = a.km;
  const int i = a.is + (int)(col % a.ni), j = a.js + (int)(col / a.ni);
  const long long cell = (long long)(j - a.jsd) * a.nid + (i - a.isd);
  CV<double> PE1{a.pe1 + col, nc}, PE2{a.pe2 + col, nc}, D1v{a.d1 + col, nc}, D2v{a.d2 + col, nc};
  column_pressures(km, a.ptop, a.ak, a.bk, a.dp1 + cell, a.s2, PE1, PE2, D1v, D2v);
  CV<D1> A1{a.a4 + col, 4 * nc}, A2{a.a4 + nc + col, 4 * nc}, A3{a.a4 + 2 * nc + col, 4 * nc},
      A4{a.a4 + 3 * nc + col, 4 * nc};
  CV<D1> Q{a.qi + col, nc}, G{a.gam + col, nc};
  CV<uint8_t> E1{a.flags + col, 3 * nc}, E5{a.flags + nc + col, 3 * nc}, E6{a.flags + 2 * nc + col, 3 * nc};
  for (int iq = 0; iq < a.nq; iq++) {
    const long long base = (long long)iq * km * a.s2 + cell;
    for (int k = 1; k <= km; k++) {
      D1 x(a.qv[base + (long long)(k - 1) * a.s2]);
      x.d[0] = a.qt[base + (long long)(k - 1) * a.s2];
      A1(k) = x;
    }
    cs_profile_col<D1>(A1, A2, A3, A4, D1v, D1(0.0), km, 0, a.kord, true, D1(0.0), Q, G, E1, E5, E6);
    double *ov = a.ov + base, *ot = a.ot + base;
    const long long s2 = a.s2;
    auto out = [=](int k, const D1 &v) {
      ov[(long long)(k - 1) * s2] = v.v;
      ot[(long long)(k - 1) * s2] = v.d[0];
    };
    map_col_s<D1>(PE1, PE2, D1v, A1, A2, A3, A4, km, a.nq > 5, D2v, out);
  }
  if (a.delp_out)
    for (int k = 1; k <= km; k++) a.delp_out[(long long)(k - 1) * a.s2 + cell] = D2v(k);
}

// Limited parabola (a2, a3, a4) of cell c from the raw interface values
// Q(c), Q(c+1), the means a1(c-2..c+2) (w[0..4]) and the primal extremum
// flags (bit 0 extm, 1 ext5, 2 ext6) of cells c-1, c, c+1: the interface
// limits and cell limiters of cs_profile / scalar_profile for iv = 0.
template <typename S>
__device__ void cell_coeffs(int c, int km, int akord, double qmin, const S (&w)[5], S Qc, S Qc1,
                            int e_lo, int e_c, int e_hi, S &a2, S &a3, S &a4) {
  const S a1 = w[2];
  if (akord > 16) {
    a2 = Qc;
    a3 = Qc1;
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
    return;
  }
  // interface limits, interface p between cells p-1 and p; w[o] = a1(c-2+o)
  auto lim = [&](int p, S q, const S &am2, const S &am1, const S &a0, const S &ap1) -> S {
    if (p == 1 || p == km + 1) return q;
    if (p == 2 || p == km) {
      q = mn2(q, mx2(am1, a0));
      return mx2(q, mn2(am1, a0));
    }
    const S glo = am1 - am2, ghi = ap1 - a0;
    if (glo * ghi > S(0.0)) {
      q = mn2(q, mx2(am1, a0));
      q = mx2(q, mn2(am1, a0));
    } else if (glo > S(0.0)) {
      q = mx2(q, mn2(am1, a0));
    } else {
      q = mn2(q, mx2(am1, a0));
      q = mx2(S(0.0), q);
    }
    return q;
  };
  const S zero(0.0);
  a2 = lim(c, Qc, c >= 3 ? w[0] : zero, c >= 2 ? w[1] : zero, w[2], c <= km - 1 ? w[3] : zero);
  a3 = lim(c + 1, Qc1, c >= 2 ? w[1] : zero, w[2], c <= km - 1 ? w[3] : zero, c <= km - 2 ? w[4] : zero);
  const bool em = (e_c & 1) != 0;
  if (c == 1) {
    a2 = mx2(S(0.0), a2);
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
    S a1c = a1;
    cs_lim1<S>(a1c, a2, a3, a4, em, 1);
    return;
  }
  if (c == km) {
    a3 = mx2(S(0.0), a3);
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
    S a1c = a1;
    cs_lim1<S>(a1c, a2, a3, a4, em, 1);
    return;
  }
  if (c == 2 || c == km - 1) {
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
    S a1c = a1;
    cs_lim1<S>(a1c, a2, a3, a4, em, 2);
    return;
  }
  const S gkm1 = w[1] - w[0], gk = w[2] - w[1], gk1 = w[3] - w[2], gk2 = w[4] - w[3];
  if (akord < 9) {
    huynh2<S>(a1, a2, a3, gk, gk1, gk2, gkm1);
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
  } else if (akord == 9) {
    if ((em && (e_lo & 1)) || (em && (e_hi & 1)) || (em && a1 < S(qmin))) {
      a2 = a1;
      a3 = a1;
      a4 = S(0.0);
    } else {
      a4 = S(6.0) * a1 - S(3.0) * (a2 + a3);
      if (fabs(a4) > fabs(a2 - a3)) {
        huynh2<S>(a1, a2, a3, gk, gk1, gk2, gkm1);
        a4 = S(6.0) * a1 - S(3.0) * (a2 + a3);
      }
    }
  } else {
    if (e_c & 2) {
      if ((e_lo & 2) || (e_hi & 2) || a1 < S(qmin)) {
        a2 = a1;
        a3 = a1;
      } else if ((e_lo & 4) || (e_hi & 4)) {
        huynh2<S>(a1, a2, a3, gk, gk1, gk2, gkm1);
      }
    } else if (e_c & 4) {
      if ((e_lo & 2) || (e_hi & 2)) huynh2<S>(a1, a2, a3, gk, gk1, gk2, gkm1);
    }
    a4 = S(3.0) * (S(2.0) * a1 - (a2 + a3));
  }
  S a1c = a1;
  cs_lim1<S>(a1c, a2, a3, a4, em, 0);
}

struct RemapAD {
  int km, nq, ni, is, js, isd, jsd, nid, kord;
  long long ncol, s2;
  const double *ak, *bk;
  double ptop;
  const double *dp1;
  const double *qv;    // transported state (remap input)
  const double *gbar;  // adjoint of the remapped tracers
  double *abar;        // adjoint of the remap input (out)
  // scratch [level][column]
  double *pe1, *pe2, *d1, *d2, *a4, *qi, *gam, *dm, *qraw, *bar;  // bar: 4 x km adjoints of a1..a4
  uint8_t *flags;
};

__global__ void __launch_bounds__(64) k_remap_ad(const RemapAD a) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  const long long nc = a.ncol;
  const int km = a.km;
  const int i = a.is + (int)(col % a.ni), j = a.js + (int)(col / a.ni);
  const long long cell = (long long)(j - a.jsd) * a.nid + (i - a.isd);
  CV<double> PE1{a.pe1 + col, nc}, PE2{a.pe2 + col, nc}, D1v{a.d1 + col, nc}, D2v{a.d2 + col, nc};
  column_pressures(km, a.ptop, a.ak, a.bk, a.dp1 + cell, a.s2, PE1, PE2, D1v, D2v);
  CV<double> A1{a.a4 + col, 4 * nc}, A2{a.a4 + nc + col, 4 * nc}, A3{a.a4 + 2 * nc + col, 4 * nc},
      A4{a.a4 + 3 * nc + col, 4 * nc};
  CV<double> Q{a.qi + col, nc}, G{a.gam + col, nc}, QR{a.qraw + col, nc};
  CV<double> B1{a.bar + col, 4 * nc}, B2{a.bar + nc + col, 4 * nc}, B3{a.bar + 2 * nc + col, 4 * nc},
      B4{a.bar + 3 * nc + col, 4 * nc};
  CV<uint8_t> E1{a.flags + col, 3 * nc}, E5{a.flags + nc + col, 3 * nc}, E6{a.flags + 2 * nc + col, 3 * nc};
  const bool mapn = a.nq > 5;
  const double r3 = 1.0 / 3.0, r23 = 2.0 / 3.0;
  for (int iq = 0; iq < a.nq; iq++) {
    const long long base = (long long)iq * km * a.s2 + cell;
    for (int k = 1; k <= km; k++) A1(k) = a.qv[base + (long long)(k - 1) * a.s2];
    // primal profile (for the extremum flags)
    cs_profile_col<double>(A1, A2, A3, A4, D1v, 0.0, km, 0, a.kord, true, 0.0, Q, G, E1, E5, E6);
    // raw interface values: the tridiagonal solve of cs_profile (iv != -2)
    double grat = D1v(2) / D1v(1);
    double bet = grat * (grat + 0.5);
    QR(1) = ((grat + grat) * (grat + 1.0) * A1(1) + A1(2)) / bet;
    G(1) = (1.0 + grat * (grat + 1.5)) / bet;
    double d4 = 0.0;
    for (int k = 2; k <= km; k++) {
      d4 = D1v(k - 1) / D1v(k);
      const double bt = 2.0 + d4 + d4 - G(k - 1);
      QR(k) = (3.0 * (A1(k - 1) + d4 * A1(k)) - QR(k - 1)) / bt;
      G(k) = d4 / bt;
    }
    const double a_bot = 1.0 + d4 * (d4 + 1.5);
    const double den_bot = d4 * (d4 + 0.5) - a_bot * G(km);
    QR(km + 1) = (2.0 * d4 * (d4 + 1.0) * A1(km) + A1(km - 1) - a_bot * QR(km)) / den_bot;
    for (int k = km; k >= 1; k--) QR(k) = QR(k) - G(k) * QR(k + 1);
    for (int k = 1; k <= km; k++) {
      B1(k) = 0.0;
      B2(k) = 0.0;
      B3(k) = 0.0;
      B4(k) = 0.0;
    }
    // ---- layer integrals, transposed (same search as the forward)
    const double *gb = a.gbar + base;
    int k0 = 1;
    for (int k = 1; k <= km; k++) {
      const double gk = gb[(long long)(k - 1) * a.s2];
      for (int l = k0; l <= km; l++) {
        if (PE2(k) >= PE1(l) && PE2(k) <= PE1(l + 1)) {
          const double pl = (PE2(k) - PE1(l)) / D1v(l);
          if (PE2(k + 1) <= PE1(l + 1)) {
            const double pr = (PE2(k + 1) - PE1(l)) / D1v(l);
            const double f1 = 0.5 * (pr + pl), f2 = r3 * (pr * (pr + pl) + pl * pl);
            B2(l) += gk * (1.0 - f1);
            B3(l) += gk * f1;
            B4(l) += gk * (f1 - f2);
            k0 = l;
            break;
          }
          const double wk = gk / D2v(k);
          {
            const double dp = PE1(l + 1) - PE2(k);
            const double f1 = 0.5 * (1.0 + pl), f2 = r3 * (1.0 + pl * (1.0 + pl));
            B2(l) += wk * dp * (1.0 - f1);
            B3(l) += wk * dp * f1;
            B4(l) += wk * dp * (f1 - f2);
          }
          for (int m = l + 1; m <= km; m++) {
            if (PE2(k + 1) > PE1(m + 1)) {
              B1(m) += wk * D1v(m);
            } else {
              const double dp = PE2(k + 1) - PE1(m);
              const double esl = dp / D1v(m);
              const double f1 = 0.5 * esl, f2 = 1.0 - r23 * esl;
              B2(m) += wk * dp * (1.0 - f1);
              B3(m) += wk * dp * f1;
              B4(m) += wk * dp * f1 * f2;
              k0 = m;
              break;
            }
          }
          break;
        }
      }
    }
    (void)mapn;
    // ---- limiters, transposed through the local Jacobian of every cell
    for (int k = 1; k <= km + 1; k++) G(k) = 0.0;  // G reused: adjoint of QR
    for (int c = 1; c <= km; c++) {
      D7 w[5];
      for (int o = 0; o < 5; o++) {
        const int kk = c - 2 + o;
        w[o] = D7(kk >= 1 && kk <= km ? A1(kk) : 0.0);
        w[o].d[o] = 1.0;
      }
      D7 qc(QR(c)), qc1(QR(c + 1));
      qc.d[5] = 1.0;
      qc1.d[6] = 1.0;
      const int e_lo = c > 1 ? (E1(c - 1) | (E5(c - 1) << 1) | (E6(c - 1) << 2)) : 0;
      const int e_c = E1(c) | (E5(c) << 1) | (E6(c) << 2);
      const int e_hi = c < km ? (E1(c + 1) | (E5(c + 1) << 1) | (E6(c + 1) << 2)) : 0;
      D7 a2, a3, a4;
      cell_coeffs<D7>(c, km, a.kord, 0.0, w, qc, qc1, e_lo, e_c, e_hi, a2, a3, a4);
      const double b2 = B2(c), b3 = B3(c), b4 = B4(c);
      for (int o = 0; o < 5; o++) {
        const int kk = c - 2 + o;
        if (kk >= 1 && kk <= km) B1(kk) += b2 * a2.d[o] + b3 * a3.d[o] + b4 * a4.d[o];
      }
      G(c) += b2 * a2.d[5] + b3 * a3.d[5] + b4 * a4.d[5];
      G(c + 1) += b2 * a2.d[6] + b3 * a3.d[6] + b4 * a4.d[6];
    }
    // ---- tridiagonal solve, transposed.  Forward: Q'(1) = r1/bet,
    // Q'(k) = (r_k - Q'(k-1))/bt_k, Qb = (r_b - a_bot Q'(km))/den_bot,
    // Q(k) = Q'(k) - g_k Q(k+1).  G now holds the adjoint of Q; recompute the
    // factors into Q (reused) as needed.
    // factors g_k and bt_k
    grat = D1v(2) / D1v(1);
    bet = grat * (grat + 0.5);
    double gprev = (1.0 + grat * (grat + 1.5)) / bet;
    Q(1) = gprev;   // g_1
    A2(1) = bet;    // bt_1 (A2 reused as bt)
    for (int k = 2; k <= km; k++) {
      d4 = D1v(k - 1) / D1v(k);
      const double bt = 2.0 + d4 + d4 - gprev;
      gprev = d4 / bt;
      Q(k) = gprev;
      A2(k) = bt;
      A3(k) = d4;  // d4_k (A3 reused)
    }
    // back substitution transposed: Q(k) = Q'(k) - g_k Q(k+1), k = km..1
    // adjoint: for k = 1..km: Q'bar(k) = Qbar(k); Qbar(k+1) -= g_k Qbar(k)
    for (int k = 1; k <= km; k++) G(k + 1) -= Q(k) * G(k);
    // bottom row: Q(km+1) = (cb a1(km) + a1(km-1) - a_bot Q'(km)) / den_bot
    const double qb = G(km + 1);
    B1(km) += 2.0 * d4 * (d4 + 1.0) / den_bot * qb;
    B1(km - 1) += qb / den_bot;
    double qpb = G(km) - a_bot / den_bot * qb;  // adjoint of Q'(km)
    for (int k = km; k >= 2; k--) {
      // Q'(k) = (3 (a1(k-1) + d4_k a1(k)) - Q'(k-1)) / bt_k
      const double t = qpb / A2(k);
      B1(k - 1) += 3.0 * t;
      B1(k) += 3.0 * A3(k) * t;
      qpb = G(k - 1) - t;  // adjoint of Q'(k-1)
    }
    // Q'(1) = (2 grat (grat + 1) a1(1) + a1(2)) / bet
    B1(1) += (grat + grat) * (grat + 1.0) / bet * qpb;
    B1(2) += qpb / bet;
    double *ab = a.abar + base;
    for (int k = 1; k <= km; k++) ab[(long long)(k - 1) * a.s2] = B1(k);
  }
}

}  // namespace

// ------------------------------------------------------------------ drivers
struct TlAdPlan {
  std::vector<int> nsplit;
  int nmax = 1;
  const int *d_nsplit = nullptr;
  const double *d_frac = nullptr;
  double *lev = nullptr;  // cx cy xfx yfx rax ray mfx mfy per level
  LevelData L{};
  double *dp_hist = nullptr;  // (nmax+1) x npz x s2
};

// Sub-cycle counts (tracer_2d / tracer_2d_1L rule) into the staging arena
// and to the host.
static int tlad_nsplit(Ctx &c, bool per_level, int q_split, const double *cx, const double *cy,
                       TlAdPlan &P) {
  const int npz = c.npz;
  hipStream_t s = c.stream;
  const size_t b8 = ((sizeof(double) * npz) + 255) & ~(size_t)255;
  char *st = (char *)ctx_stage(c, 2 * b8 + sizeof(int) * npz + 256);
  if (!st) return -1;
  double *cmax = (double *)st, *frac = (double *)(st + b8);
  int *nspl = (int *)(st + 2 * b8);
  if (device_nsplit<double>(c, cx, cy, per_level, q_split, cmax, nspl, frac)) return -1;
  P.nsplit.assign(npz, 1);
  FV3_HIP(hipMemcpyAsync(P.nsplit.data(), nspl, sizeof(int) * npz, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  P.nmax = 1;
  for (int k = 0; k < npz; k++) {
    if (P.nsplit[k] < 1) return set_error("nsplit(%d)=%d from a non-finite Courant number", k + 1, P.nsplit[k]);
    P.nmax = std::max(P.nmax, P.nsplit[k]);
  }
  P.d_nsplit = nspl;
  P.d_frac = frac;
  return 0;
}

// Scaled winds, area fluxes and the thickness of every sub-cycle.
static int tlad_prepare(Ctx &c, int t, const double *cx, const double *cy, const double *mfx,
                        const double *mfy, const double *dp1, char *&cursor, TlAdPlan &P) {
  const Dims d = c.d;
  const int npz = c.npz;
  hipStream_t s = c.stream;
  const GridDev<double> g = grid_of<double>(c);
  // tile t's inputs and metrics
  cx += (long long)t * d.sx * npz;
  cy += (long long)t * d.sy * npz;
  mfx += (long long)t * d.smx * npz;
  mfy += (long long)t * d.smy * npz;
  dp1 += (long long)t * d.s2 * npz;
  const long long tm = (long long)t * g.s2;
  auto take = [&](size_t bytes) {
    char *r = cursor;
    cursor += (bytes + 255) & ~(size_t)255;
    return r;
  };
  double *cxs = (double *)take(sizeof(double) * d.sx * npz), *cys = (double *)take(sizeof(double) * d.sy * npz);
  double *xfx = (double *)take(sizeof(double) * d.sx * npz), *yfx = (double *)take(sizeof(double) * d.sy * npz);
  double *rax = (double *)take(sizeof(double) * d.srx * npz), *ray = (double *)take(sizeof(double) * d.sry * npz);
  double *mfxs = (double *)take(sizeof(double) * d.smx * npz), *mfys = (double *)take(sizeof(double) * d.smy * npz);
  const int *nspl = P.d_nsplit;
  const double *frac = P.d_frac;
  const double *dxa = g.dxa + tm, *dya = g.dya + tm, *gdx = g.dx + (long long)t * g.sdx,
               *gdy = g.dy + (long long)t * g.sdy, *sg = g.sin_sg + 9 * tm, *area = g.area + tm;
  const long long gs2 = g.s2;
  const int *nsp = nspl;
  const double *fr = frac;
  // area fluxes on the upwind metric, then the 1/nsplit scaling
  if (box(s, [=] __device__(int i, int j, int k) {
        const long long lv = (long long)k * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
        const long long m0 = (long long)(j - d.jsd) * d.nid;
        const double cv = cx[lv];
        const double dyv = gdy[(long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)];
        double x;
        if (cv > 0.0)
          x = cv * dxa[m0 + (i - 1 - d.isd)] * dyv * sg[2 * gs2 + m0 + (i - 1 - d.isd)];
        else
          x = cv * dxa[m0 + (i - d.isd)] * dyv * sg[0 * gs2 + m0 + (i - d.isd)];
        const bool sc = nsp[k] > 1;
        cxs[lv] = sc ? cv * fr[k] : cv;
        xfx[lv] = sc ? x * fr[k] : x;
      }, d.is, d.ie + 1, d.jsd, d.jed, npz))
    return -1;
  if (box(s, [=] __device__(int i, int j, int k) {
        const long long lv = (long long)k * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
        const double cv = cy[lv];
        const double dxv = gdx[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
        double y;
        if (cv > 0.0)
          y = cv * dya[(long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)] * dxv * sg[3 * gs2 + (long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)];
        else
          y = cv * dya[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * dxv * sg[1 * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
        const bool sc = nsp[k] > 1;
        cys[lv] = sc ? cv * fr[k] : cv;
        yfx[lv] = sc ? y * fr[k] : y;
      }, d.isd, d.ied, d.js, d.je + 1, npz))
    return -1;
  if (box(s, [=] __device__(int i, int j, int k) {
        const long long o = (long long)k * d.smx + (long long)(j - d.js) * d.nix + (i - d.is);
        mfxs[o] = nsp[k] > 1 ? mfx[o] * fr[k] : mfx[o];
      }, d.is, d.ie + 1, d.js, d.je, npz))
    return -1;
  if (box(s, [=] __device__(int i, int j, int k) {
        const long long o = (long long)k * d.smy + (long long)(j - d.js) * d.niy + (i - d.is);
        mfys[o] = nsp[k] > 1 ? mfy[o] * fr[k] : mfy[o];
      }, d.is, d.ie, d.js, d.je + 1, npz))
    return -1;
  if (box(s, [=] __device__(int i, int j, int k) {
        const double *xk = xfx + (long long)k * d.sx + (long long)(j - d.jsd) * d.nix;
        rax[(long long)k * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
            area[(long long)(j - d.jsd) * d.nid + (i - d.isd)] + xk[i - d.is] - xk[i + 1 - d.is];
      }, d.is, d.ie, d.jsd, d.jed, npz))
    return -1;
  if (box(s, [=] __device__(int i, int j, int k) {
        const double *yk = yfx + (long long)k * d.sy;
        ray[(long long)k * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
            area[(long long)(j - d.jsd) * d.nid + (i - d.isd)] + yk[(long long)(j - d.js) * d.nid + (i - d.isd)] -
            yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
      }, d.isd, d.ied, d.js, d.je, npz))
    return -1;
  P.L = LevelData{cxs, cys, xfx, yfx, rax, ray, mfxs, mfys};
  // thickness before each sub-cycle: dp(0) = dp1, dp(it) = dp(it-1) + div(mf) rarea
  // on levels with nsplit >= it (the running dp1 <- dp2 of tracer_2d)
  P.dp_hist = (double *)take(sizeof(double) * d.s2 * npz * (P.nmax + 1));
  FV3_HIP(hipMemcpyAsync(P.dp_hist, dp1, sizeof(double) * d.s2 * npz, hipMemcpyDeviceToDevice, s));
  const double *rarea = g.rarea + tm;
  for (int it = 1; it <= P.nmax; it++) {
    const double *prev = P.dp_hist + (long long)(it - 1) * npz * d.s2;
    double *cur = P.dp_hist + (long long)it * npz * d.s2;
    FV3_HIP(hipMemcpyAsync(cur, prev, sizeof(double) * d.s2 * npz, hipMemcpyDeviceToDevice, s));
    const int itc = it;
    if (box(s, [=] __device__(int i, int j, int k) {
          if (nsp[k] < itc) return;
          const double *mx = mfxs + (long long)k * d.smx + (long long)(j - d.js) * d.nix;
          const double *my = mfys + (long long)k * d.smy;
          const long long o = (long long)k * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd);
          cur[o] = prev[o] + ((mx[i - d.is] - mx[i + 1 - d.is]) +
                              (my[(long long)(j - d.js) * d.niy + (i - d.is)] - my[(long long)(j + 1 - d.js) * d.niy + (i - d.is)])) *
                                 rarea[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
        }, d.is, d.ie, d.js, d.je, npz))
      return -1;
  }
  return 0;
}

static size_t tlad_prep_bytes(const Dims &d, int npz, int nmax) {
  size_t b = 0;
  auto add = [&](size_t x) { b += (x + 255) & ~(size_t)255; };
  add(sizeof(double) * d.sx * npz); add(sizeof(double) * d.sy * npz);
  add(sizeof(double) * d.sx * npz); add(sizeof(double) * d.sy * npz);
  add(sizeof(double) * d.srx * npz); add(sizeof(double) * d.sry * npz);
  add(sizeof(double) * d.smx * npz); add(sizeof(double) * d.smy * npz);
  add(sizeof(double) * d.s2 * npz * (nmax + 1));
  return b;
}

static int tlad_check(Ctx &c, int hord, int kord_tr) {
  if (c.ntiles != 1 && !(c.ntiles == 6 && c.halo_n > 0))
    return set_error("tangent-linear/adjoint: a single-tile or six-face context is required");
  if (hord < 7 || hord > 13) return set_error("hord=%d: only the monotone orders 7..13 are implemented", hord);
  if (kord_tr != 0 && !(kord_tr > 7 && (kord_tr <= 10 || kord_tr > 16)))
    return set_error("kord_tr=%d: implemented orders are 8, 9, 10 and > 16 (0 skips the remap)", kord_tr);
  if (kord_tr != 0 && c.npz < 6) return set_error("remap tangent-linear/adjoint needs npz >= 6");
  return 0;
}

// The halo table grouped by source cell, for k_halo_adjoint (built once per
// context from the forward table).
static int halo_adjoint_table(Ctx &c, const long long **usrc, const int **ptr, const long long **dst, int *nsrc) {
  if (!c.halo_adj) {
    const int n = c.halo_n;
    std::vector<long long> hd(n), hs(n);
    FV3_HIP(hipMemcpy(hd.data(), c.halo_dst, sizeof(long long) * n, hipMemcpyDeviceToHost));
    FV3_HIP(hipMemcpy(hs.data(), c.halo_src, sizeof(long long) * n, hipMemcpyDeviceToHost));
    std::vector<int> ord(n);
    for (int e = 0; e < n; e++) ord[e] = e;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return hs[x] < hs[y]; });
    std::vector<long long> us, ds;
    std::vector<int> pt;
    for (int k = 0; k < n; k++) {
      const int e = ord[k];
      if (us.empty() || us.back() != hs[e]) {
        us.push_back(hs[e]);
        pt.push_back(k);
      }
      ds.push_back(hd[e]);
    }
    pt.push_back(n);
    const int nu = (int)us.size();
    const size_t b_us = ((sizeof(long long) * nu) + 255) & ~(size_t)255;
    const size_t b_pt = ((sizeof(int) * (nu + 1)) + 255) & ~(size_t)255;
    void *buf = nullptr;
    FV3_HIP(hipMalloc(&buf, b_us + b_pt + sizeof(long long) * std::max(n, 1)));
    char *cb = (char *)buf;
    FV3_HIP(hipMemcpy(cb, us.data(), sizeof(long long) * nu, hipMemcpyHostToDevice));
    FV3_HIP(hipMemcpy(cb + b_us, pt.data(), sizeof(int) * (nu + 1), hipMemcpyHostToDevice));
    FV3_HIP(hipMemcpy(cb + b_us + b_pt, ds.data(), sizeof(long long) * n, hipMemcpyHostToDevice));
    c.halo_adj = buf;
    c.halo_adj_nsrc = nu;
  }
  const int nu = c.halo_adj_nsrc;
  const size_t b_us = ((sizeof(long long) * nu) + 255) & ~(size_t)255;
  const size_t b_pt = ((sizeof(int) * (nu + 1)) + 255) & ~(size_t)255;
  char *cb = (char *)c.halo_adj;
  *usrc = (const long long *)cb;
  *ptr = (const int *)(cb + b_us);
  *dst = (const long long *)(cb + b_us + b_pt);
  *nsrc = nu;
  return 0;
}

// per-tile metric views of the context's FP64 grid
static Geo geo_of(const Ctx &c, int t) {
  const GridDev<double> &g = c.grid64;
  const long long tm = (long long)t * g.s2;
  return Geo{c.d, g.area + tm, g.rarea + tm, g.dxa + tm, g.dya + tm, c.npz};
}

int step_tl(Ctx &c, bool per_level, int q_split, int nq, const double *q0, const double *dq0,
            double *q1, double *dq1, const double *dp1, double *delp_out, const double *cx,
            const double *cy, const double *mfx, const double *mfy, const double *ak,
            const double *bk, double ptop, int hord, int kord_tr) {
  if (tlad_check(c, hord, kord_tr)) return 1;
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles;
  hipStream_t s = c.stream;
  const long long qt = d.s2 * npz * nq, qsize = qt * nt;  // one tile, all tiles
  TlAdPlan P0;
  if (tlad_nsplit(c, per_level, q_split, cx, cy, P0)) return 1;
  const int nb = P0.nmax;
  const int nslab = npz * nq;
  const size_t per_slab = sizeof(D1) * (d.sy + d.sry + d.smx + d.sx + d.srx + d.smy);
  const int chunk = (int)std::max<long long>(1, std::min<long long>(std::min(nslab, 65535), (2ll << 30) / (long long)per_slab));
  const long long ncol = (long long)(d.ie - d.is + 1) * (d.je - d.js + 1);
  const size_t remap_b = kord_tr ? sizeof(double) * ncol * (4 * (npz + 1)) + sizeof(D1) * ncol * (7 * (npz + 1)) +
                                       3 * ncol * (npz + 1) + 4096
                                 : 0;
  const size_t bytes = tlad_prep_bytes(d, npz, nb) * nt + 2 * sizeof(double) * qsize + sizeof(int) * nslab +
                       per_slab * chunk + remap_b + 8192 * nt;
  char *base = (char *)ctx_scratch(c, bytes);
  if (!base) return 1;
  char *cur = base;
  std::vector<TlAdPlan> P(nt, P0);
  for (int t = 0; t < nt; t++)
    if (tlad_prepare(c, t, cx, cy, mfx, mfy, dp1, cur, P[t])) return 1;
  auto take = [&](size_t b) {
    char *r = cur;
    cur += (b + 255) & ~(size_t)255;
    return r;
  };
  double *wv = (double *)take(sizeof(double) * qsize), *wt = (double *)take(sizeof(double) * qsize);
  int *zs = (int *)take(sizeof(int) * nslab);
  D1 *tmp = (D1 *)take(per_slab * chunk);
  // working state: q0 everywhere, tangent dq0 on the interior only
  FV3_HIP(hipMemcpyAsync(wv, q0, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  if (box(s, [=] __device__(int i, int j, int z) {
        const long long o = (long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd);
        const bool in = i >= d.is && i <= d.ie && j >= d.js && j <= d.je;
        wt[o] = in ? dq0[o] : 0.0;
      }, d.isd, d.ied, d.jsd, d.jed, nslab * nt))
    return 1;
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  std::vector<int> act;
  for (int it = 1; it <= nb; it++) {
    act.clear();
    for (int iq = 0; iq < nq; iq++)
      for (int k = 0; k < npz; k++)
        if (P0.nsplit[k] >= it) act.push_back(iq * npz + k);
    FV3_HIP(hipMemcpyAsync(zs, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
    // six faces: every sub-cycle starts from halos filled from the
    // neighbouring faces' current interiors, values and tangents alike
    if (nt > 1 && (halo_fill<double>(c, wv, qt, nullptr, 0, nslab) || halo_fill<double>(c, wt, qt, nullptr, 0, nslab)))
      return 1;
    for (int t = 0; t < nt; t++) {
      const Geo G = geo_of(c, t);
      const double *dpa = P[t].dp_hist + (long long)(it - 1) * npz * d.s2, *dpb = P[t].dp_hist + (long long)it * npz * d.s2;
      for (size_t p0 = 0; p0 < act.size(); p0 += chunk) {
        const int np = (int)std::min<size_t>(chunk, act.size() - p0);
        SubTemps<D1> T{tmp, tmp + d.sy * chunk, tmp + (d.sy + d.sry) * chunk,
                       tmp + (d.sy + d.sry + d.smx) * chunk, tmp + (d.sy + d.sry + d.smx + d.sx) * chunk,
                       tmp + (d.sy + d.sry + d.smx + d.sx + d.srx) * chunk};
        if (subcycle_forward<D1>(s, G, P[t].L, zs + p0, np, wv + t * qt, wt + t * qt, T, dpa, dpb, ord_in, ord_ou,
                                 true))
          return 1;
      }
    }
    FV3_HIP(hipStreamSynchronize(s));  // act is reused by the next sub-cycle
  }
  // outputs (halo cells keep the transported values; the remap rewrites the interior)
  FV3_HIP(hipMemcpyAsync(q1, wv, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  FV3_HIP(hipMemcpyAsync(dq1, wt, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  if (delp_out) FV3_HIP(hipMemcpyAsync(delp_out, dp1, sizeof(double) * d.s2 * npz * nt, hipMemcpyDeviceToDevice, s));
  if (!kord_tr) {
    FV3_HIP(hipStreamSynchronize(s));
    return 0;
  }
  RemapTL r{};
  r.km = npz; r.nq = nq; r.ni = d.ie - d.is + 1; r.is = d.is; r.js = d.js; r.isd = d.isd; r.jsd = d.jsd;
  r.nid = d.nid; r.kord = kord_tr; r.ncol = ncol; r.s2 = d.s2; r.ak = ak; r.bk = bk; r.ptop = ptop;
  r.pe1 = (double *)take(sizeof(double) * ncol * (npz + 1));
  r.pe2 = (double *)take(sizeof(double) * ncol * (npz + 1));
  r.d1 = (double *)take(sizeof(double) * ncol * (npz + 1));
  r.d2 = (double *)take(sizeof(double) * ncol * (npz + 1));
  r.a4 = (D1 *)take(sizeof(D1) * ncol * 4 * npz);
  r.qi = (D1 *)take(sizeof(D1) * ncol * (npz + 1));
  r.gam = (D1 *)take(sizeof(D1) * ncol * (npz + 1));
  r.dm = nullptr;
  r.flags = (uint8_t *)take(3 * ncol * npz);
  if ((size_t)(cur - base) > bytes) return set_error("tangent-linear scratch layout overflow");
  for (int t = 0; t < nt; t++) {  // column-local: one tile after another on the same scratch
    const long long td = (long long)t * d.s2 * npz;
    r.dp1 = dp1 + td;
    r.delp_out = delp_out ? delp_out + td : nullptr;
    r.qv = wv + t * qt; r.qt = wt + t * qt; r.ov = q1 + t * qt; r.ot = dq1 + t * qt;
    hipLaunchKernelGGL(k_remap_tl, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, s, r);
    FV3_HIP(hipGetLastError());
  }
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

int step_ad(Ctx &c, bool per_level, int q_split, int nq, const double *q0, const double *qbar1,
            double *qbar0, const double *dp1, const double *cx, const double *cy,
            const double *mfx, const double *mfy, const double *ak, const double *bk,
            double ptop, int hord, int kord_tr) {
  if (tlad_check(c, hord, kord_tr)) return 1;
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles;
  const bool faces = nt > 1;
  hipStream_t s = c.stream;
  const long long qt = d.s2 * npz * nq, qsize = qt * nt;
  TlAdPlan P0;
  if (tlad_nsplit(c, per_level, q_split, cx, cy, P0)) return 1;
  const int nb = P0.nmax;
  const int nslab = npz * nq;
  const long long ncol = (long long)(d.ie - d.is + 1) * (d.je - d.js + 1);
  // per slab: primal temps, 4 x 6 weights, adjoint temps, corner gathers
  const long long faces_n = d.sy + d.smx + d.sx + d.smy;
  const size_t per_slab = sizeof(double) * (d.sy + d.sry + d.smx + d.sx + d.srx + d.smy) +
                          sizeof(double) * 6 * faces_n +
                          sizeof(double) * (d.s2 + d.smx + d.smy + d.smx + d.sx + d.smy + d.sy + d.srx + d.sry + 36);
  const int chunk = (int)std::max<long long>(1, std::min<long long>(std::min(nslab, 65535), (2ll << 30) / (long long)per_slab));
  const size_t remap_b = kord_tr ? sizeof(double) * ncol * (npz + 1) * 16 + 3 * ncol * (npz + 1) + 4096 : 0;
  // checkpoints: the state entering every sub-cycle, plus the transported state
  const size_t bytes = tlad_prep_bytes(d, npz, nb) * nt + sizeof(double) * qsize * (nb + 2) + sizeof(int) * nslab +
                       per_slab * chunk + remap_b + 8192 * nt;
  const long long *hu = nullptr, *hd = nullptr;
  const int *hp = nullptr;
  int hn = 0;
  if (faces && halo_adjoint_table(c, &hu, &hp, &hd, &hn)) return 1;
  char *base = (char *)ctx_scratch(c, bytes);
  if (!base) return 1;
  char *cur = base;
  std::vector<TlAdPlan> P(nt, P0);
  for (int t = 0; t < nt; t++)
    if (tlad_prepare(c, t, cx, cy, mfx, mfy, dp1, cur, P[t])) return 1;
  auto take = [&](size_t b) {
    char *r = cur;
    cur += (b + 255) & ~(size_t)255;
    return r;
  };
  double *ck = (double *)take(sizeof(double) * qsize * (nb + 1));  // ck[it-1]: input of sub-cycle it; ck[nmax]: final
  double *qbar = (double *)take(sizeof(double) * qsize);
  int *zs = (int *)take(sizeof(int) * nslab);
  char *tmp = take(per_slab * chunk);
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  // temporaries
  double *tp = (double *)tmp;
  SubTemps<double> T{tp, tp + d.sy * chunk, tp + (d.sy + d.sry) * chunk, tp + (d.sy + d.sry + d.smx) * chunk,
                     tp + (d.sy + d.sry + d.smx + d.sx) * chunk, tp + (d.sy + d.sry + d.smx + d.sx + d.srx) * chunk};
  double *wp = tp + (d.sy + d.sry + d.smx + d.sx + d.srx + d.smy) * chunk;
  SweepW W{wp, wp + 6 * d.sy * chunk, wp + 6 * (d.sy + d.smx) * chunk, wp + 6 * (d.sy + d.smx + d.sx) * chunk};
  double *ap = wp + 6 * faces_n * chunk;
  AdjTemps A;
  A.qa = ap;
  A.Fx = A.qa + d.s2 * chunk;
  A.Fy = A.Fx + d.smx * chunk;
  A.fx = A.Fy + d.smy * chunk;
  A.fx2 = A.fx + d.smx * chunk;
  A.fy = A.fx2 + d.sx * chunk;
  A.fy2 = A.fy + d.smy * chunk;
  A.qj = A.fy2 + d.sy * chunk;
  A.qi = A.qj + d.srx * chunk;
  A.cg = A.qi + d.sry * chunk;
  // ---- forward sweep with checkpoints (six faces: halos filled before every sub-cycle)
  FV3_HIP(hipMemcpyAsync(ck, q0, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  std::vector<std::vector<int>> acts(nb + 1);
  for (int it = 1; it <= nb; it++) {
    for (int iq = 0; iq < nq; iq++)
      for (int k = 0; k < npz; k++)
        if (P0.nsplit[k] >= it) acts[it].push_back(iq * npz + k);
  }
  for (int it = 1; it <= nb; it++) {
    double *prev = ck + (long long)(it - 1) * qsize, *next = ck + (long long)it * qsize;
    if (faces && halo_fill<double>(c, prev, qt, nullptr, 0, nslab)) return 1;
    FV3_HIP(hipMemcpyAsync(next, prev, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
    const std::vector<int> &act = acts[it];
    FV3_HIP(hipMemcpyAsync(zs, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
    for (int t = 0; t < nt; t++) {
      const Geo G = geo_of(c, t);
      const double *dpa = P[t].dp_hist + (long long)(it - 1) * npz * d.s2, *dpb = P[t].dp_hist + (long long)it * npz * d.s2;
      for (size_t p0 = 0; p0 < act.size(); p0 += chunk) {
        const int np = (int)std::min<size_t>(chunk, act.size() - p0);
        if (subcycle_forward<double>(s, G, P[t].L, zs + p0, np, next + t * qt, nullptr, T, dpa, dpb, ord_in, ord_ou,
                                     true))
          return 1;
      }
    }
    FV3_HIP(hipStreamSynchronize(s));
  }
  const double *qT = ck + (long long)nb * qsize;
  // ---- remap adjoint: qbar = (d remap / d qT)^T qbar1 on the interior
  FV3_HIP(hipMemsetAsync(qbar, 0, sizeof(double) * qsize, s));
  if (kord_tr) {
    RemapAD r{};
    r.km = npz; r.nq = nq; r.ni = d.ie - d.is + 1; r.is = d.is; r.js = d.js; r.isd = d.isd; r.jsd = d.jsd;
    r.nid = d.nid; r.kord = kord_tr; r.ncol = ncol; r.s2 = d.s2; r.ak = ak; r.bk = bk; r.ptop = ptop;
    r.pe1 = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.pe2 = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.d1 = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.d2 = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.a4 = (double *)take(sizeof(double) * ncol * 4 * npz);
    r.qi = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.gam = (double *)take(sizeof(double) * ncol * (npz + 2));
    r.dm = nullptr;
    r.qraw = (double *)take(sizeof(double) * ncol * (npz + 1));
    r.bar = (double *)take(sizeof(double) * ncol * 4 * npz);
    r.flags = (uint8_t *)take(3 * ncol * npz);
    if ((size_t)(cur - base) > bytes) return set_error("adjoint scratch layout overflow");
    for (int t = 0; t < nt; t++) {
      r.dp1 = dp1 + (long long)t * d.s2 * npz;
      r.qv = qT + t * qt; r.gbar = qbar1 + t * qt; r.abar = qbar + t * qt;
      hipLaunchKernelGGL(k_remap_ad, dim3((unsigned)((ncol + 63) / 64)), dim3(64), 0, s, r);
      FV3_HIP(hipGetLastError());
    }
  } else {
    if ((size_t)(cur - base) > bytes) return set_error("adjoint scratch layout overflow");
    FV3_HIP(hipMemcpyAsync(qbar, qbar1, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  }
  // the adjoint of the step's output lives on the interior only
  auto clear_halo = [&]() {
    return box(s, [=] __device__(int i, int j, int z) {
          const bool in = i >= d.is && i <= d.ie && j >= d.js && j <= d.je;
          if (!in) qbar[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] = 0.0;
        }, d.isd, d.ied, d.jsd, d.jed, nslab * nt);
  };
  if (clear_halo()) return 1;
  // ---- reverse sweep over sub-cycles, recomputing each from its checkpoint
  for (int it = nb; it >= 1; it--) {
    const std::vector<int> &act = acts[it];
    FV3_HIP(hipMemcpyAsync(zs, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice, s));
    double *qin = ck + (long long)(it - 1) * qsize;
    for (int t = 0; t < nt; t++) {
      const Geo G = geo_of(c, t);
      const double *dpa = P[t].dp_hist + (long long)(it - 1) * npz * d.s2, *dpb = P[t].dp_hist + (long long)it * npz * d.s2;
      for (size_t p0 = 0; p0 < act.size(); p0 += chunk) {
        const int np = (int)std::min<size_t>(chunk, act.size() - p0);
        if (subcycle_forward<double>(s, G, P[t].L, zs + p0, np, qin + t * qt, nullptr, T, dpa, dpb, ord_in, ord_ou,
                                     false))
          return 1;
        if (sweep_weights(s, G, P[t].L, zs + p0, np, qin + t * qt, T, W, ord_in, ord_ou)) return 1;
        if (subcycle_adjoint(s, G, P[t].L, zs + p0, np, qbar + t * qt, W, A, dpa, dpb, faces)) return 1;
      }
    }
    if (faces) {  // transposed halo fill: halo adjoints onto the neighbours' interiors
      for (int s0 = 0; s0 < nslab; s0 += 65535) {
        const int ns = std::min(65535, nslab - s0);
        hipLaunchKernelGGL(k_halo_adjoint, dim3((unsigned)((hn + 255) / 256), (unsigned)ns), dim3(256), 0, s,
                           qbar + (long long)s0 * d.s2, qt, d.s2, hu, hp, hd, hn);
        FV3_HIP(hipGetLastError());
      }
      if (clear_halo()) return 1;
    }
    FV3_HIP(hipStreamSynchronize(s));
  }
  FV3_HIP(hipMemcpyAsync(qbar0, qbar, sizeof(double) * qsize, hipMemcpyDeviceToDevice, s));
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace fv3
