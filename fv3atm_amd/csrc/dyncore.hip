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
// FV3_FORM_PLAIN64 runs every (tile, level) slab of a sub-step in one launch
// per stage (the staged transport's stage kernels, in the FP64 oracle's
// operation order); FV3_FORM_FUSED64 / FUSED32 run a whole sub-step in one
// launch of blocks that keep every intermediate in LDS, with the same
// expressions, so both FP64 forms match the oracle bit for bit.
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


// ---- one acoustic sub-step in one launch ---------------------------------------
// A workgroup owns a TX x TY block of one (tile, level) slab and computes every
// stage of the sub-step for it in LDS, reading delp (with its 3-cell halo),
// the sub-step's winds and the metrics once: the Courant numbers and area
// fluxes, the inner y-sweep of delp and q_i, the outer x-sweep, the inner
// x-sweep and q_j, the outer y-sweep, the flux averages, the accumulations
// and the delp update.  Each quantity is the same expression as in the stage
// kernels above, so the FP64 form rounds as they do.  delp is read from one
// buffer and written to another (a neighbouring block still reads the old
// values as its halo); the halo cells are carried over, and the corner-block
// cells get their x-sweep (dir 1) copy_corners sources, as the stages leave
// them in place.
struct DynSubArgs {
  Dims d;
  int npz, ord_in, ord_ou, first;
  long long gs2, gdx, gdy;
};

template <typename T, int TX, int TY>
struct DynTile {
  static constexpr int C = TX + 6, R = TY + 6;  // delp block with its halo: columns i0-3 .. i1+3
  static constexpr int FX = TX + 1;             // x-faces i0 .. i1+1
  static constexpr int FY = TY + 1;             // y-faces j0 .. j1+1
  T py[R * C], px[R * C];                       // delp with the y- (dir 2) and x-sweep (dir 1) corners
  T crx[R * FX], xfx[R * FX], fx2[R * FX];      // x-faces of rows j0-3 .. j1+3
  T cry[FY * C], yfx[FY * C];                   // y-faces of columns i0-3 .. i1+3
  T fy2[FY * C];                                // y-faces of columns i0-3 .. i1+3
  T qi[TY * C];                                 // rows j0 .. j1, columns i0-3 .. i1+3
  T fx[TY * FX];                                // x-faces of rows j0 .. j1
  T qj[R * TX];                                 // rows j0-3 .. j1+3, columns i0 .. i1
  T fy[FY * TX];                                // y-faces of columns i0 .. i1
};

template <typename T>
struct LdsLine {  // cells of one LDS line indexed by their global index
  const T *p;
  int lo, stride;
  __device__ __forceinline__ T operator()(int m) const { return p[(m - lo) * stride]; }
};

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(256) k_dyn_sub(const DynSubArgs a, const GridDev<T> g, T bdt, const T *u,
                                                 const T *v, const T *din, T *dout, T *cx, T *cy, T *mfx, T *mfy) {
  using TL = DynTile<T, TX, TY>;
  constexpr int C = TL::C, R = TL::R, FX = TL::FX, FY = TL::FY;
  extern __shared__ __align__(16) unsigned char dyn_lds[];
  TL &L = *reinterpret_cast<TL *>(dyn_lds);
  const Dims &d = a.d;
  const int tid = threadIdx.x, z = blockIdx.z, tile = z / a.npz;
  const int i0 = d.is + (int)blockIdx.x * TX, j0 = d.js + (int)blockIdx.y * TY;
  const int i1 = min(i0 + TX - 1, d.ie), j1 = min(j0 + TY - 1, d.je);
  const T *dl = din + (long long)z * d.s2;
  const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
  const long long gt = tile * a.gs2;
  auto cell = [&](int i, int j) { return (long long)(j - d.jsd) * d.nid + (i - d.isd); };
  auto in_dom = [&](int i, int j) { return i >= d.isd && i <= d.ied && j >= d.jsd && j <= d.jed; };
  // only blocks whose halo reaches a cube corner have copy_corners cells
  const bool corner_blk = !d.bounded && (i0 - 3 <= 0 || i1 + 3 >= d.npx) && (j0 - 3 <= 0 || j1 + 3 >= d.npy);

  // ---- delp block (both corner versions)
  for (int e = tid; e < R * C; e += 256) {
    const int r = e / C, c = e % C, i = i0 - 3 + c, j = j0 - 3 + r;
    T vy = T(0), vx = T(0);
    if (in_dom(i, j)) {
      vy = vx = dl[cell(i, j)];
      int si, sj;
      if (corner_blk) {
        if (corner_src(i, j, d.npx, d.npy, 2, d.corners, si, sj)) vy = dl[cell(si, sj)];
        if (corner_src(i, j, d.npx, d.npy, 1, d.corners, si, sj)) vx = dl[cell(si, sj)];
      }
    }
    L.py[e] = vy;
    L.px[e] = vx;
  }
  // ---- Courant numbers and area fluxes with the upwind metric; cx, cy accumulate
  for (int e = tid; e < R * FX; e += 256) {
    const int r = e / FX, f = e % FX, i = i0 + f, j = j0 - 3 + r;
    if (i > d.ie + 1 || j < d.jsd || j > d.jed) continue;
    const long long lv = (long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
    // both upwind candidates, so that no address depends on the loaded wind
    const T w = u[lv];
    const T da_l = g.dxa[gt + cell(i - 1, j)], da_r = g.dxa[gt + cell(i, j)];
    const T sg_l = g.sin_sg[tile * 9 * a.gs2 + 2 * a.gs2 + cell(i - 1, j)];
    const T sg_r = g.sin_sg[tile * 9 * a.gs2 + cell(i, j)];
    const T dyf = g.dy[tile * a.gdy + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)];
    const bool pos = w > T(0);
    const T dau = pos ? da_l : da_r;
    const T cr = bdt * w / dau;
    L.crx[e] = cr;
    L.xfx[e] = cr * dau * dyf * (pos ? sg_l : sg_r);
    // rows and faces this block owns: its own, the halo rows at the domain edges, face ie+1
    const bool own_j = (j >= j0 && j <= j1) || (j < d.js && j0 == d.js) || (j > d.je && j1 == d.je);
    const bool own_i = i <= i1 || (i == d.ie + 1 && i1 == d.ie);
    if (own_j && own_i) cx[lv] = a.first ? T(0) + cr : cx[lv] + cr;
  }
  for (int e = tid; e < FY * C; e += 256) {
    const int f = e / C, c = e % C, i = i0 - 3 + c, j = j0 + f;
    if (j > d.je + 1 || i < d.isd || i > d.ied) continue;
    const long long lv = (long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
    const T w = v[lv];
    const T da_l = g.dya[gt + cell(i, j - 1)], da_r = g.dya[gt + cell(i, j)];
    const T sg_l = g.sin_sg[tile * 9 * a.gs2 + 3 * a.gs2 + cell(i, j - 1)];
    const T sg_r = g.sin_sg[tile * 9 * a.gs2 + 1 * a.gs2 + cell(i, j)];
    const T dxf = g.dx[tile * a.gdx + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
    const bool pos = w > T(0);
    const T dau = pos ? da_l : da_r;
    const T cr = bdt * w / dau;
    L.cry[e] = cr;
    L.yfx[e] = cr * dau * dxf * (pos ? sg_l : sg_r);
    const bool own_i = (i >= i0 && i <= i1) || (i < d.is && i0 == d.is) || (i > d.ie && i1 == d.ie);
    const bool own_j = j <= j1 || (j == d.je + 1 && j1 == d.je);
    if (own_i && own_j) cy[lv] = a.first ? T(0) + cr : cy[lv] + cr;
  }
  __syncthreads();
  // ---- inner sweeps: y over delp (dir 2 corners) at faces j0 .. j1+1 of
  // columns i0-3 .. i1+3, x over delp (dir 1) at faces i0 .. i1+1 of rows j0-3 .. j1+3
  for (int e = tid; e < FY * C; e += 256) {
    const int f = e / C, c = e % C, i = i0 - 3 + c, j = j0 + f;
    if (j > j1 + 1 || i < d.isd || i > d.ied) continue;
    const LdsLine<T> qa{L.py + c, j0 - 3, C};
    const ColQ<T> da{g.dya + gt + (i - d.isd), d.jsd, d.nid};
    L.fy2[e] = ppm_flux<T>(qa, da, j, L.cry[e], gy, a.ord_in);
  }
  for (int e = tid; e < R * FX; e += 256) {
    const int r = e / FX, f = e % FX, i = i0 + f, j = j0 - 3 + r;
    if (i > i1 + 1 || j < d.jsd || j > d.jed) continue;
    const LdsLine<T> qa{L.px + r * C, i0 - 3, 1};
    const RowQ<T> da{g.dxa + gt + (long long)(j - d.jsd) * d.nid, d.isd};
    L.fx2[e] = ppm_flux<T>(qa, da, i, L.crx[e], gx, a.ord_in);
  }
  __syncthreads();
  // ---- q_i at rows j0 .. j1, columns i0-3 .. i1+3; q_j at rows j0-3 .. j1+3, columns i0 .. i1
  for (int e = tid; e < TY * C; e += 256) {
    const int r = e / C, c = e % C, i = i0 - 3 + c, j = j0 + r;
    if (j > j1 || i < d.isd || i > d.ied) continue;
    const T ya = L.yfx[r * C + c], yb = L.yfx[(r + 1) * C + c];
    const T fa = ya * L.fy2[r * C + c], fb = yb * L.fy2[(r + 1) * C + c];
    const T ar = g.area[gt + cell(i, j)];
    L.qi[e] = (L.py[(r + 3) * C + c] * ar + fa - fb) / (ar + ya - yb);
  }
  for (int e = tid; e < R * TX; e += 256) {
    const int r = e / TX, c = e % TX, i = i0 + c, j = j0 - 3 + r;
    if (i > i1 || j < d.jsd || j > d.jed) continue;
    const T xa = L.xfx[r * FX + c], xb = L.xfx[r * FX + c + 1];
    const T fa = xa * L.fx2[r * FX + c], fb = xb * L.fx2[r * FX + c + 1];
    const T ar = g.area[gt + cell(i, j)];
    L.qj[e] = (L.px[r * C + c + 3] * ar + fa - fb) / (ar + xa - xb);
  }
  __syncthreads();
  // ---- outer sweeps (x over q_i, y over q_j), averaged with the inner ones
  // and the area fluxes; mfx, mfy accumulate
  for (int e = tid; e < TY * FX; e += 256) {
    const int r = e / FX, f = e % FX, i = i0 + f, j = j0 + r;
    if (i > i1 + 1 || j > j1) continue;
    const LdsLine<T> qa{L.qi + r * C, i0 - 3, 1};
    const RowQ<T> da{g.dxa + gt + (long long)(j - d.jsd) * d.nid, d.isd};
    const T fo = ppm_flux<T>(qa, da, i, L.crx[(r + 3) * FX + f], gx, a.ord_ou);
    const T fl = T(0.5) * (fo + L.fx2[(r + 3) * FX + f]) * L.xfx[(r + 3) * FX + f];
    L.fx[e] = fl;
    if (i <= i1 || i == d.ie + 1) {
      const long long o = (long long)z * d.smx + (long long)(j - d.js) * d.nix + (i - d.is);
      mfx[o] = a.first ? T(0) + fl : mfx[o] + fl;
    }
  }
  for (int e = tid; e < FY * TX; e += 256) {
    const int f = e / TX, c = e % TX, i = i0 + c, j = j0 + f;
    if (i > i1 || j > j1 + 1) continue;
    const LdsLine<T> qa{L.qj + c, j0 - 3, TX};
    const ColQ<T> da{g.dya + gt + (i - d.isd), d.jsd, d.nid};
    const T fo = ppm_flux<T>(qa, da, j, L.cry[f * C + c + 3], gy, a.ord_ou);
    const T fl = T(0.5) * (fo + L.fy2[f * C + c + 3]) * L.yfx[f * C + c + 3];
    L.fy[e] = fl;
    if (j <= j1 || j == d.je + 1) {
      const long long o = (long long)z * d.smy + (long long)(j - d.js) * d.niy + (i - d.is);
      mfy[o] = a.first ? T(0) + fl : mfy[o] + fl;
    }
  }
  __syncthreads();
  // ---- delp: the block's cells by the flux divergence; its share of the halo
  // carried over (corner blocks with their dir 1 sources)
  T *dw = dout + (long long)z * d.s2;
  for (int e = tid; e < R * C; e += 256) {
    const int r = e / C, c = e % C, i = i0 - 3 + c, j = j0 - 3 + r;
    if (!in_dom(i, j)) continue;
    const bool inner = i >= i0 && i <= i1 && j >= j0 && j <= j1;
    if (inner) {
      const int ri = j - j0, ci = i - i0;
      dw[cell(i, j)] = L.py[e] + ((L.fx[ri * FX + ci] - L.fx[ri * FX + ci + 1]) +
                                  (L.fy[ri * TX + ci] - L.fy[(ri + 1) * TX + ci])) *
                                     g.rarea[gt + cell(i, j)];
    } else {
      // halo cells of the slab, each written by the one block whose range
      // (extended to the slab edge on the sides where it touches it) holds it
      const bool hi_ok = (i >= i0 && i <= i1) || (i < d.is && i0 == d.is) || (i > d.ie && i1 == d.ie);
      const bool hj_ok = (j >= j0 && j <= j1) || (j < d.js && j0 == d.js) || (j > d.je && j1 == d.je);
      const bool halo = i < d.is || i > d.ie || j < d.js || j > d.je;
      if (halo && hi_ok && hj_ok) dw[cell(i, j)] = L.px[e];
    }
  }
}

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

// The sub-steps of dyn_core_fluxes, one launch each (plus the six-face halo
// refresh between them); delp alternates between the caller's array and a
// scratch copy and ends in the caller's.
template <typename T>
static int dyn_core_fluxes_fused(Ctx &c, int n_split, T bdt, int hord, const T *uc, const T *vc, T *delp, T *dp1,
                                 T *cx, T *cy, T *mfx, T *mfy, hipStream_t s) {
  constexpr int TX = 32, TY = sizeof(T) == 8 ? 8 : 16;
  using TL = DynTile<T, TX, TY>;
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles, nz = npz * nt;
  const GridDev<T> g = grid_of<T>(c);
  const long long slab = (long long)d.s2 * nz;
  T *alt = (T *)ctx_scratch(c, sizeof(T) * slab);
  if (!alt) return 1;
  FV3_HIP(hipMemcpyAsync(dp1, delp, sizeof(T) * slab, hipMemcpyDeviceToDevice, s));
  DynSubArgs a{};
  a.d = d;
  a.npz = npz;
  a.ord_in = hord == 10 ? 8 : hord;
  a.ord_ou = hord;
  a.gs2 = g.s2;
  a.gdx = g.sdx;
  a.gdy = g.sdy;
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  const dim3 grid((unsigned)((ni + TX - 1) / TX), (unsigned)((nj + TY - 1) / TY), (unsigned)nz);
  T *buf[2] = {delp, alt};
  for (int it = 0; it < n_split; it++) {
    a.first = it == 0;
    const T *u = uc + (long long)it * nz * d.sx, *v = vc + (long long)it * nz * d.sy;
    T *din = buf[it & 1], *dout = buf[(it + 1) & 1];
    hipLaunchKernelGGL((k_dyn_sub<T, TX, TY>), grid, dim3(256), sizeof(TL), s, a, g, bdt, u, v, din, dout, cx, cy,
                       mfx, mfy);
    FV3_HIP(hipGetLastError());
    if (nt > 1 && it + 1 < n_split)
      if (halo_fill<T>(c, dout, (long long)npz * d.s2, nullptr, 0, npz)) return 1;
  }
  if (n_split & 1) FV3_HIP(hipMemcpyAsync(delp, alt, sizeof(T) * slab, hipMemcpyDeviceToDevice, s));
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
  else if (form == FV3_FORM_FUSED64)
    rc = dyn_core_fluxes_fused<double>(c, n_split, bdt, hord_dp, (const double *)uc, (const double *)vc,
                                       (double *)delp, (double *)dp1, (double *)cx, (double *)cy, (double *)mfx,
                                       (double *)mfy, s);
  else if (form == FV3_FORM_FUSED32)
    rc = dyn_core_fluxes_fused<float>(c, n_split, (float)bdt, hord_dp, (const float *)uc, (const float *)vc,
                                      (float *)delp, (float *)dp1, (float *)cx, (float *)cy, (float *)mfx,
                                      (float *)mfy, s);
  else
    rc = set_error("form %d: the dyn_core fluxes run FV3_FORM_PLAIN64 or FV3_FORM_FUSED64 (double), "
                   "FV3_FORM_FUSED32 (float)", form);
  if (rc) return 1;
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}
