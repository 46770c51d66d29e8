// tp_plain.hip -- the plain (staged) form of tracer_2d / tracer_2d_1L /
// fv_tp_2d (upstream fv_tracer2d.F90, tp_core.F90) on MI355X.
//
// Each stage of fv_tp_2d is its own kernel writing its result to HBM, exactly
// as the Fortran writes its local arrays: corner fill, inner y-sweep (fy2),
// q_i, outer x-sweep (fx), corner fill, inner x-sweep (fx2), q_j, outer
// y-sweep (fy), flux averaging, [deln_flux], flux-form update.  Launches are
// batched over every active (tile, level, tracer) slab of a sub-cycle, one
// slab per grid.z index.  This form is the FP64 parity form (bit-compatible
// with an FP64 CPU evaluation) and the baseline the fused FP32 form is
// measured against.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "fused.h"
#include "fv3_internal.h"

namespace fv3 {

struct Slab {
  int tile, k, iq;
};

template <typename F>
__global__ void __launch_bounds__(256) k_rect(F f, int i0, int i1, int j0, int j1) {
  const int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int j = j0 + (int)(blockIdx.y * blockDim.y + threadIdx.y);
  if (i <= i1 && j <= j1) f(i, j, (int)blockIdx.z);
}

template <typename F>
static int launch_rect(hipStream_t s, F f, int i0, int i1, int j0, int j1, int nz) {
  if (i1 < i0 || j1 < j0 || nz <= 0) return 0;
  dim3 b(64, 4);
  dim3 g((unsigned)((i1 - i0 + 64) / 64), (unsigned)((j1 - j0 + 4) / 4), (unsigned)nz);
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

template <typename T>
__device__ __forceinline__ void atomic_max_nonneg(T *addr, T v);
template <>
__device__ __forceinline__ void atomic_max_nonneg<double>(double *addr, double v) {
  atomicMax((unsigned long long *)addr, (unsigned long long)__double_as_longlong(v));
}
template <>
__device__ __forceinline__ void atomic_max_nonneg<float>(float *addr, float v) {
  atomicMax((unsigned int *)addr, (unsigned int)__float_as_int(v));
}

// per-level cmax: block reduction then one atomic per block (max is order
// free).  Evaluated in FP64 from the FP64 sin_sg for every form, so that
// the FP32 form's sub-cycle counts differ from FP64's only through the
// caller's own FP32 rounding of the Courant numbers.
template <typename T>
__global__ void __launch_bounds__(256) k_cmax(const T *cx, const T *cy, const double *sin_sg, Dims d,
                                              int npz, long long gs2, double *cmax) {
  const int z = blockIdx.z, tile = z / npz, k = z % npz;
  const T *cxk = cx + ((long long)tile * npz + k) * d.sx;
  const T *cyk = cy + ((long long)tile * npz + k) * d.sy;
  const double *s5 = sin_sg + (long long)tile * 9 * gs2 + 4 * gs2;
  const bool top = (k + 1) < npz / 6;
  double m = 0.0;
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < ni * nj; t += gridDim.x * blockDim.x) {
    const int i = d.is + t % ni, j = d.js + t / ni;
    const double ax = fabs((double)cxk[(long long)(j - d.jsd) * d.nix + (i - d.is)]);
    const double ay = fabs((double)cyk[(long long)(j - d.js) * d.nid + (i - d.isd)]);
    if (top)
      m = mx3(m, ax, ay);
    else
      m = mx2(m, mx2(ax, ay) + 1.0 - s5[(long long)(j - d.jsd) * d.nid + (i - d.isd)]);
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = mx2(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomic_max_nonneg(&cmax[k], red[0]);
}

// nsplit(k) and 1/nsplit from cmax (tracer_2d_1L per level, tracer_2d global)
template <typename T>
__global__ void k_nsplit(const double *cmax, int npz, int per_level, int q_split, int *nsplit,
                         T *frac) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int glob = 1;
  if (!per_level) {
    if (q_split == 0) {
      double c_global = cmax[0];
      if (npz != 1)
        for (int k = 1; k < npz; k++) c_global = mx2(cmax[k], c_global);
      glob = (int)(1.0 + c_global);
    } else {
      glob = q_split;
    }
  }
  for (int k = 0; k < npz; k++) {
    int n = per_level ? (int)(1.0 + cmax[k]) : glob;
    nsplit[k] = n;
    frac[k] = T(1) / T(n);
  }
}

// zero fill by a kernel: unlike a memset node it is safe inside a captured graph
template <typename T>
__global__ void k_zero(T *p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = T(0);
}

template <typename T>
int device_nsplit(Ctx &c, const T *cx, const T *cy, bool per_level, int q_split, double *cmax,
                  int *nsplit, T *frac) {
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles;
  hipLaunchKernelGGL(k_zero<double>, dim3(1), dim3(256), 0, c.stream, cmax, npz);
  FV3_HIP(hipGetLastError());
  dim3 b(256), gr((unsigned)std::min(64, (d.ie - d.is + 1) * (d.je - d.js + 1) / 256 + 1), 1,
                  (unsigned)(npz * nt));
  hipLaunchKernelGGL(k_cmax<T>, gr, b, 0, c.stream, cx, cy, c.grid64.sin_sg, d, npz, c.grid64.s2, cmax);
  FV3_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_nsplit<T>, dim3(1), dim3(1), 0, c.stream, cmax, npz, per_level ? 1 : 0,
                     q_split, nsplit, frac);
  FV3_HIP(hipGetLastError());
  return 0;
}
template int device_nsplit<double>(Ctx &, const double *, const double *, bool, int, double *, int *, double *);
template int device_nsplit<float>(Ctx &, const float *, const float *, bool, int, double *, int *, float *);

template <typename T>
static size_t align_up(size_t b) {
  return (b + 255) & ~(size_t)255;
}

template <typename T>
int tracer_2d_plain(Ctx &c, bool per_level, T *q, T *dp1, T *mfx, T *mfy, T *cx, T *cy, int nq,
                    int hord, int q_split, T dt, int id_divg, int nord_tr, T trdm, T lim_fac,
                    int *nsplit_host) {
  (void)lim_fac;
  const Dims d = c.d;
  const int npz = c.npz, nt = c.ntiles;
  const GridDev<T> g = grid_of<T>(c);
  hipStream_t s = c.stream;
  if (hord < 7 || hord > 13) return set_error("hord=%d: only the monotone orders 7..13 are implemented", hord);
  if (nord_tr < 0 || nord_tr > 2) return set_error("nord_tr=%d: deln_flux supports nord 0..2 with a 3-cell halo", nord_tr);
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  const bool damp_on = !per_level && trdm > T(1.0e-4);

  // ---- scratch layout
  const long long tq = d.s2 * npz * nq;  // tile stride of q
  size_t off = 0;
  auto take = [&](size_t bytes) {
    size_t o = off;
    off += align_up<T>(bytes);
    return o;
  };
  const size_t o_xfx = take(sizeof(T) * d.sx * npz * nt);
  const size_t o_yfx = take(sizeof(T) * d.sy * npz * nt);
  const size_t o_rax = take(sizeof(T) * d.srx * npz * nt);
  const size_t o_ray = take(sizeof(T) * d.sry * npz * nt);
  const size_t o_dp2 = take(sizeof(T) * d.sdp * npz * nt);
  const size_t o_cmax = take(sizeof(double) * npz);
  const size_t o_frac = take(sizeof(T) * npz);
  const size_t o_nspl = take(sizeof(int) * npz);
  const size_t o_lev = take(sizeof(int) * npz * nt);
  const size_t o_qn2 = per_level ? take(sizeof(T) * tq * nt) : take(0);
  const long long per_slab = d.sy + d.sry + d.smx + d.sx + d.srx + d.smy +
                             (damp_on ? d.s2 + (long long)(d.nid + 1) * d.njd + (long long)d.nid * (d.njd + 1) : 0);
  const long long nslab_all = (long long)nt * npz * nq;
  const long long budget = 3ll << 30;
  int chunk = (int)std::max(1ll, std::min(nslab_all, budget / (long long)(sizeof(T) * per_slab)));
  chunk = std::min(chunk, 8192);
  const size_t o_pairs = take(sizeof(Slab) * chunk);
  const size_t o_slab = take(sizeof(int) * npz * nq);
  const size_t o_tmp = take(sizeof(T) * per_slab * chunk);
  char *base = (char *)ctx_scratch(c, off);
  if (!base) return -1;
  T *xfx = (T *)(base + o_xfx), *yfx = (T *)(base + o_yfx);
  T *rax = (T *)(base + o_rax), *ray = (T *)(base + o_ray), *dp2 = (T *)(base + o_dp2);
  double *cmax = (double *)(base + o_cmax);
  T *fracd = (T *)(base + o_frac);
  int *nspl = (int *)(base + o_nspl), *levs = (int *)(base + o_lev);
  T *qn2 = (T *)(base + o_qn2);
  Slab *pairs = (Slab *)(base + o_pairs);
  int *slab_dev = (int *)(base + o_slab);
  T *tmp = (T *)(base + o_tmp);
  T *t_fy2 = tmp, *t_qi = t_fy2 + d.sy * chunk, *t_fx = t_qi + d.sry * chunk;
  T *t_fx2 = t_fx + d.smx * chunk, *t_qj = t_fx2 + d.sx * chunk, *t_fy = t_qj + d.srx * chunk;
  T *t_d2 = t_fy + d.smy * chunk;
  T *t_dfx = t_d2 + d.s2 * chunk;
  T *t_dfy = t_dfx + (long long)(d.nid + 1) * d.njd * chunk;

  // ---- area fluxes (upwind metric), per level
  {
    const long long gs2 = g.s2, gdy = g.sdy, gdx = g.sdx;
    const T *dxa = g.dxa, *dya = g.dya, *gdxp = g.dx, *gdyp = g.dy, *sg = g.sin_sg;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const long long lv = (long long)z * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is);
          const long long m0 = tile * gs2 + (long long)(j - d.jsd) * d.nid;
          const T cv = cx[lv];
          const T dyv = gdyp[tile * gdy + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)];
          if (cv > T(0))
            xfx[lv] = cv * dxa[m0 + (i - 1 - d.isd)] * dyv * sg[tile * 9 * gs2 + 2 * gs2 + (long long)(j - d.jsd) * d.nid + (i - 1 - d.isd)];
          else
            xfx[lv] = cv * dxa[m0 + (i - d.isd)] * dyv * sg[tile * 9 * gs2 + 0 * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
        }, d.is, d.ie + 1, d.jsd, d.jed, npz * nt))
      return -1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const long long lv = (long long)z * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd);
          const T cv = cy[lv];
          const T dxv = gdxp[tile * gdx + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
          if (cv > T(0))
            yfx[lv] = cv * dya[tile * gs2 + (long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)] * dxv *
                      sg[tile * 9 * gs2 + 3 * gs2 + (long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)];
          else
            yfx[lv] = cv * dya[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] * dxv *
                      sg[tile * 9 * gs2 + 1 * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
        }, d.isd, d.ied, d.js, d.je + 1, npz * nt))
      return -1;
  }
  // ---- cmax, nsplit, in-place scaling
  FV3_HIP(hipMemsetAsync(cmax, 0, sizeof(double) * npz, s));
  {
    dim3 b(256), gr((unsigned)std::min(64, (d.ie - d.is + 1) * (d.je - d.js + 1) / 256 + 1), 1,
                    (unsigned)(npz * nt));
    hipLaunchKernelGGL(k_cmax<T>, gr, b, 0, s, cx, cy, c.grid64.sin_sg, d, npz, c.grid64.s2, cmax);
    FV3_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_nsplit<T>, dim3(1), dim3(1), 0, s, cmax, npz, per_level ? 1 : 0,
                       q_split, nspl, fracd);
    FV3_HIP(hipGetLastError());
  }
  std::vector<int> nsplit(npz);
  FV3_HIP(hipMemcpyAsync(nsplit.data(), nspl, sizeof(int) * npz, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  int nmax = 1;
  for (int k = 0; k < npz; k++) {
    if (nsplit[k] < 1) return set_error("nsplit(%d)=%d from a non-finite Courant number", k + 1, nsplit[k]);
    nmax = std::max(nmax, nsplit[k]);
  }
  if (nsplit_host)
    for (int k = 0; k < npz; k++) nsplit_host[k] = nsplit[k];
  {
    const int *nsp = nspl;
    const T *fr = fracd;
    auto scale2 = [=] __device__(T * a, long long sz, int i, int j, int z, int nx, int jlo) {
      (void)sz;
      const int k = z % npz;
      if (nsp[k] > 1) {
        T *p = a + (long long)z * sz + (long long)(j - jlo) * nx + i;
        *p = *p * fr[k];
      }
    };
    // cx, xfx (is:ie+1, jsd:jed)
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          scale2(cx, d.sx, i - d.is, j, z, d.nix, d.jsd);
          scale2(xfx, d.sx, i - d.is, j, z, d.nix, d.jsd);
        }, d.is, d.ie + 1, d.jsd, d.jed, npz * nt))
      return -1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) { scale2(mfx, d.smx, i - d.is, j, z, d.nix, d.js); },
                    d.is, d.ie + 1, d.js, d.je, npz * nt))
      return -1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          scale2(cy, d.sy, i - d.isd, j, z, d.nid, d.js);
          scale2(yfx, d.sy, i - d.isd, j, z, d.nid, d.js);
        }, d.isd, d.ied, d.js, d.je + 1, npz * nt))
      return -1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) { scale2(mfy, d.smy, i - d.is, j, z, d.niy, d.js); },
                    d.is, d.ie, d.js, d.je + 1, npz * nt))
      return -1;
  }
  // ---- ra_x, ra_y per level
  {
    const T *area = g.area;
    const long long gs2 = g.s2;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *xk = xfx + (long long)z * d.sx + (long long)(j - d.jsd) * d.nix;
          rax[(long long)z * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
              area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + xk[i - d.is] - xk[i + 1 - d.is];
        }, d.is, d.ie, d.jsd, d.jed, npz * nt))
      return -1;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *yk = yfx + (long long)z * d.sy;
          ray[(long long)z * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
              area[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] +
              yk[(long long)(j - d.js) * d.nid + (i - d.isd)] - yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
        }, d.isd, d.ied, d.js, d.je, npz * nt))
      return -1;
  }

  if (nt > 1 && per_level) {  // complete_group_halo_update(q_pack)
    if (halo_fill<T>(c, q, tq, nullptr, 0, npz * nq)) return -1;
  }

  std::vector<Slab> hp;
  std::vector<int> hl;
  for (int it = 1; it <= nmax; it++) {
    if (nt > 1 && !per_level) {
      if (halo_fill<T>(c, q, tq, nullptr, 0, npz * nq)) return -1;
    }
    hl.clear();
    for (int t = 0; t < nt; t++)
      for (int k = 0; k < npz; k++)
        if (nsplit[k] >= it) hl.push_back(t * npz + k);
    const int nlev = (int)hl.size();
    FV3_HIP(hipMemcpyAsync(levs, hl.data(), sizeof(int) * nlev, hipMemcpyHostToDevice, s));
    // dp2 = dp1 + div(mf) * rarea on active levels
    {
      const int *lv = levs;
      const T *rarea = g.rarea;
      const long long gs2 = g.s2;
      if (launch_rect(s, [=] __device__(int i, int j, int zz) {
            const int z = lv[zz], tile = z / npz;
            const T *mx = mfx + (long long)z * d.smx + (long long)(j - d.js) * d.nix;
            const T *my = mfy + (long long)z * d.smy;
            dp2[(long long)z * d.sdp + (long long)(j - d.js) * d.niy + (i - d.is)] =
                dp1[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] +
                ((mx[i - d.is] - mx[i + 1 - d.is]) +
                 (my[(long long)(j - d.js) * d.niy + (i - d.is)] - my[(long long)(j + 1 - d.js) * d.niy + (i - d.is)])) *
                    rarea[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
          }, d.is, d.ie, d.js, d.je, nlev))
        return -1;
    }
    // qn2 = q at the first sub-cycle of split levels
    if (per_level && it == 1) {
      const int *lv = levs;
      const int *nsp = nspl;
      if (launch_rect(s, [=] __device__(int i, int j, int zz) {
            const int iq = zz % nq, z = lv[zz / nq], tile = z / npz, k = z % npz;
            if (nsp[k] > 1) {
              const long long o = tile * tq + ((long long)iq * npz + k) * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd);
              qn2[o] = q[o];
            }
          }, d.isd, d.ied, d.jsd, d.jed, nlev * nq))
        return -1;
    }
    // all active slabs, in chunks
    hp.clear();
    for (int zz = 0; zz < nlev; zz++)
      for (int iq = 0; iq < nq; iq++) hp.push_back(Slab{hl[zz] / npz, hl[zz] % npz, iq});
    for (size_t p0 = 0; p0 < hp.size(); p0 += chunk) {
      const int np = (int)std::min<size_t>(chunk, hp.size() - p0);
      FV3_HIP(hipMemcpyAsync(pairs, hp.data() + p0, sizeof(Slab) * np, hipMemcpyHostToDevice, s));
      const Slab *pr = pairs;
      const int *nsp = nspl;
      const long long gs2 = g.s2;
      const T *area = g.area, *dxa = g.dxa, *dya = g.dya, *rarea = g.rarea;
      const bool pl = per_level;
      const int itc = it;
      auto work_of = [=] __device__(const Slab &sl) -> T * {
        const long long o = sl.tile * tq + ((long long)sl.iq * npz + sl.k) * d.s2;
        return (pl && nsp[sl.k] > 1) ? qn2 + o : q + o;
      };
      const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
      const unsigned corners = d.corners;
      // copy_corners(dir=2)
      if (!d.bounded)
        if (launch_rect(s, [=] __device__(int cidx, int unused, int p) {
              (void)unused;
              const Slab sl = pr[p];
              T *w = work_of(sl);
              const int blk = cidx / 9, cc = cidx % 9;
              const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + cc % 3;
              const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + cc / 3;
              int si, sj;
              if (corner_src(i, j, d.npx, d.npy, 2, corners, si, sj))
                w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = w[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
            }, 0, 35, 0, 0, np))
          return -1;
      // fy2 = yppm(q, cry, ord_in) over (isd:ied, js:je+1)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const T *w = work_of(sl);
            ColQ<T> qa{w + (i - d.isd), d.jsd, d.nid};
            ColQ<T> da{dya + sl.tile * gs2 + (i - d.isd), d.jsd, d.nid};
            const T cv = cy[((long long)sl.tile * npz + sl.k) * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)];
            t_fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)] = ppm_flux<T>(qa, da, j, cv, gy, ord_in);
          }, d.isd, d.ied, d.js, d.je + 1, np))
        return -1;
      // q_i = (q*area + fyy(j) - fyy(j+1)) / ra_y over (isd:ied, js:je)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const T *w = work_of(sl);
            const long long z = (long long)sl.tile * npz + sl.k;
            const T *yk = yfx + z * d.sy;
            const T *f2 = t_fy2 + (long long)p * d.sy;
            const T fa = yk[(long long)(j - d.js) * d.nid + (i - d.isd)] * f2[(long long)(j - d.js) * d.nid + (i - d.isd)];
            const T fb = yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)] * f2[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
            t_qi[(long long)p * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)] =
                (w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * area[sl.tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
                ray[z * d.sry + (long long)(j - d.js) * d.nid + (i - d.isd)];
          }, d.isd, d.ied, d.js, d.je, np))
        return -1;
      // fx = xppm(q_i, crx, ord_ou) over (is:ie+1, js:je)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            RowQ<T> qa{t_qi + (long long)p * d.sry + (long long)(j - d.js) * d.nid, d.isd};
            RowQ<T> da{dxa + sl.tile * gs2 + (long long)(j - d.jsd) * d.nid, d.isd};
            const T cv = cx[((long long)sl.tile * npz + sl.k) * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)];
            t_fx[(long long)p * d.smx + (long long)(j - d.js) * d.nix + (i - d.is)] = ppm_flux<T>(qa, da, i, cv, gx, ord_ou);
          }, d.is, d.ie + 1, d.js, d.je, np))
        return -1;
      // copy_corners(dir=1)
      if (!d.bounded)
        if (launch_rect(s, [=] __device__(int cidx, int unused, int p) {
              (void)unused;
              const Slab sl = pr[p];
              T *w = work_of(sl);
              const int blk = cidx / 9, cc = cidx % 9;
              const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + cc % 3;
              const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + cc / 3;
              int si, sj;
              if (corner_src(i, j, d.npx, d.npy, 1, corners, si, sj))
                w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = w[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
            }, 0, 35, 0, 0, np))
          return -1;
      // fx2 = xppm(q, crx, ord_in) over (is:ie+1, jsd:jed)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const T *w = work_of(sl);
            RowQ<T> qa{w + (long long)(j - d.jsd) * d.nid, d.isd};
            RowQ<T> da{dxa + sl.tile * gs2 + (long long)(j - d.jsd) * d.nid, d.isd};
            const T cv = cx[((long long)sl.tile * npz + sl.k) * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)];
            t_fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)] = ppm_flux<T>(qa, da, i, cv, gx, ord_in);
          }, d.is, d.ie + 1, d.jsd, d.jed, np))
        return -1;
      // q_j = (q*area + xfx*fx2|i - xfx*fx2|i+1) / ra_x over (is:ie, jsd:jed)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const T *w = work_of(sl);
            const long long z = (long long)sl.tile * npz + sl.k;
            const T *xk = xfx + z * d.sx + (long long)(j - d.jsd) * d.nix;
            const T *f2 = t_fx2 + (long long)p * d.sx + (long long)(j - d.jsd) * d.nix;
            const T fa = xk[i - d.is] * f2[i - d.is];
            const T fb = xk[i + 1 - d.is] * f2[i + 1 - d.is];
            t_qj[(long long)p * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)] =
                (w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * area[sl.tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
                rax[z * d.srx + (long long)(j - d.jsd) * d.niy + (i - d.is)];
          }, d.is, d.ie, d.jsd, d.jed, np))
        return -1;
      // fy = yppm(q_j, cry, ord_ou) over (is:ie, js:je+1)
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            ColQ<T> qa{t_qj + (long long)p * d.srx + (i - d.is), d.jsd, d.niy};
            ColQ<T> da{dya + sl.tile * gs2 + (i - d.isd), d.jsd, d.nid};
            const T cv = cy[((long long)sl.tile * npz + sl.k) * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)];
            t_fy[(long long)p * d.smy + (long long)(j - d.js) * d.niy + (i - d.is)] = ppm_flux<T>(qa, da, j, cv, gy, ord_ou);
          }, d.is, d.ie, d.js, d.je + 1, np))
        return -1;
      // flux averaging with the mass fluxes
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const long long z = (long long)sl.tile * npz + sl.k;
            if (i <= d.ie + 1 && j <= d.je) {
              const long long o = (long long)(j - d.js) * d.nix + (i - d.is);
              T &f = t_fx[(long long)p * d.smx + o];
              f = T(0.5) * (f + t_fx2[(long long)p * d.sx + (long long)(j - d.jsd) * d.nix + (i - d.is)]) * mfx[z * d.smx + o];
            }
            if (i <= d.ie && j <= d.je + 1) {
              const long long o = (long long)(j - d.js) * d.niy + (i - d.is);
              T &f = t_fy[(long long)p * d.smy + o];
              f = T(0.5) * (f + t_fy2[(long long)p * d.sy + (long long)(j - d.js) * d.nid + (i - d.isd)]) * mfy[z * d.smy + o];
            }
          }, d.is, d.ie + 1, d.js, d.je + 1, np))
        return -1;
      // deln_flux on the first sub-cycle of tracer_2d when trdm > 1e-4
      if (damp_on && it == 1) {
        const T damp = (T)pow((double)trdm * c.da_min, (double)(nord_tr + 1));
        const int nord = nord_tr;
        const long long nfx2 = (long long)(d.nid + 1) * d.njd, nfy2 = (long long)d.nid * (d.njd + 1);
        const T *d6u = g.del6_u, *d6v = g.del6_v;
        const long long gdx = g.sdx, gdy = g.sdy;
        // d2 = q (mass-weighted form); the untouched rim is zero
        if (launch_rect(s, [=] __device__(int i, int j, int p) {
              const Slab sl = pr[p];
              const T *w = work_of(sl);
              const bool in = i >= d.is - 1 - nord && i <= d.ie + 1 + nord && j >= d.js - 1 - nord && j <= d.je + 1 + nord;
              t_d2[(long long)p * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] =
                  in ? w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] : T(0);
            }, d.isd, d.ied, d.jsd, d.jed, np))
          return -1;
        auto corners_d2 = [&](int dir) {
          return launch_rect(s, [=] __device__(int cidx, int unused, int p) {
                (void)unused;
                T *w = t_d2 + (long long)p * d.s2;
                const int blk = cidx / 9, cc = cidx % 9;
                const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + cc % 3;
                const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + cc / 3;
                int si, sj;
                if (corner_src(i, j, d.npx, d.npy, dir, corners, si, sj))
                  w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = w[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
              }, 0, 35, 0, 0, np);
        };
        auto fx2_pass = [&](int lo, int sign) {  // dfx2 over (is-lo:ie+lo+1, js-lo:je+lo)
          return launch_rect(s, [=] __device__(int i, int j, int p) {
                const Slab sl = pr[p];
                const T *w = t_d2 + (long long)p * d.s2 + (long long)(j - d.jsd) * d.nid;
                const T dv = sign > 0 ? (w[i - 1 - d.isd] - w[i - d.isd]) : (w[i - d.isd] - w[i - 1 - d.isd]);
                t_dfx[(long long)p * nfx2 + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)] =
                    d6v[sl.tile * gdy + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)] * dv;
              }, d.is - lo, d.ie + lo + 1, d.js - lo, d.je + lo, np);
        };
        auto fy2_pass = [&](int lo, int sign) {
          return launch_rect(s, [=] __device__(int i, int j, int p) {
                const Slab sl = pr[p];
                const T *w = t_d2 + (long long)p * d.s2;
                const T a = w[(long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)], b = w[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
                const T dv = sign > 0 ? (a - b) : (b - a);
                t_dfy[(long long)p * nfy2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] =
                    d6u[sl.tile * gdx + (long long)(j - d.jsd) * d.nid + (i - d.isd)] * dv;
              }, d.is - lo, d.ie + lo, d.js - lo, d.je + lo + 1, np);
        };
        if (nord > 0 && corners_d2(1)) return -1;
        if (fx2_pass(nord, 1)) return -1;
        if (nord > 0 && corners_d2(2)) return -1;
        if (fy2_pass(nord, 1)) return -1;
        for (int n = 1; n <= nord; n++) {
          const int ntt = nord - n;
          if (launch_rect(s, [=] __device__(int i, int j, int p) {
                const Slab sl = pr[p];
                const T *fxx = t_dfx + (long long)p * nfx2 + (long long)(j - d.jsd) * (d.nid + 1);
                const T *fyy = t_dfy + (long long)p * nfy2;
                t_d2[(long long)p * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] =
                    (fxx[i - d.isd] - fxx[i + 1 - d.isd] + fyy[(long long)(j - d.jsd) * d.nid + (i - d.isd)] -
                     fyy[(long long)(j + 1 - d.jsd) * d.nid + (i - d.isd)]) *
                    rarea[sl.tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
              }, d.is - ntt - 1, d.ie + ntt + 1, d.js - ntt - 1, d.je + ntt + 1, np))
            return -1;
          if (corners_d2(1)) return -1;
          if (fx2_pass(ntt, -1)) return -1;
          if (corners_d2(2)) return -1;
          if (fy2_pass(ntt, -1)) return -1;
        }
        const T damp2 = T(0.5) * damp;
        if (launch_rect(s, [=] __device__(int i, int j, int p) {
              const Slab sl = pr[p];
              const T *mass = dp1 + ((long long)sl.tile * npz + sl.k) * d.s2;
              if (i <= d.ie + 1 && j <= d.je) {
                T &f = t_fx[(long long)p * d.smx + (long long)(j - d.js) * d.nix + (i - d.is)];
                f = f + damp2 * (mass[(long long)(j - d.jsd) * d.nid + (i - 1 - d.isd)] + mass[(long long)(j - d.jsd) * d.nid + (i - d.isd)]) *
                            t_dfx[(long long)p * nfx2 + (long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)];
              }
              if (i <= d.ie && j <= d.je + 1) {
                T &f = t_fy[(long long)p * d.smy + (long long)(j - d.js) * d.niy + (i - d.is)];
                f = f + damp2 * (mass[(long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)] + mass[(long long)(j - d.jsd) * d.nid + (i - d.isd)]) *
                            t_dfy[(long long)p * nfy2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
              }
            }, d.is, d.ie + 1, d.js, d.je + 1, np))
          return -1;
      }
      // flux-form update
      if (launch_rect(s, [=] __device__(int i, int j, int p) {
            const Slab sl = pr[p];
            const long long z = (long long)sl.tile * npz + sl.k;
            const long long o = sl.tile * tq + ((long long)sl.iq * npz + sl.k) * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd);
            const T *w = work_of(sl);
            const T *fxr = t_fx + (long long)p * d.smx + (long long)(j - d.js) * d.nix;
            const T *fyc = t_fy + (long long)p * d.smy;
            const T val = (w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * dp1[z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] +
                           (fxr[i - d.is] - fxr[i + 1 - d.is] + fyc[(long long)(j - d.js) * d.niy + (i - d.is)] -
                            fyc[(long long)(j + 1 - d.js) * d.niy + (i - d.is)]) *
                               rarea[sl.tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)]) /
                          dp2[z * d.sdp + (long long)(j - d.js) * d.niy + (i - d.is)];
            T *out = (pl && itc < nsp[sl.k]) ? qn2 : q;
            out[o] = val;
          }, d.is, d.ie, d.js, d.je, np))
        return -1;
    }
    // dp1 <- dp2 on levels that sub-cycle again
    {
      const int *lv = levs;
      const int *nsp = nspl;
      const int itc = it;
      if (launch_rect(s, [=] __device__(int i, int j, int zz) {
            const int z = lv[zz];
            if (itc < nsp[z % npz])
              dp1[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] =
                  dp2[(long long)z * d.sdp + (long long)(j - d.js) * d.niy + (i - d.is)];
          }, d.is, d.ie, d.js, d.je, nlev))
        return -1;
    }
    if (nt > 1 && per_level && it < nmax) {  // mpp_update_domains(qn2)
      std::vector<int> slabs;
      for (int k = 0; k < npz; k++)
        if (it < nsplit[k])
          for (int iq = 0; iq < nq; iq++) slabs.push_back(iq * npz + k);
      if (!slabs.empty()) {
        FV3_HIP(hipMemcpyAsync(slab_dev, slabs.data(), sizeof(int) * slabs.size(), hipMemcpyHostToDevice, s));
        if (halo_fill<T>(c, qn2, tq, slab_dev, (int)slabs.size(), 0)) return -1;
      }
    }
    FV3_HIP(hipStreamSynchronize(s));  // host vectors reused next sub-cycle
  }

  if (!per_level && id_divg > 0) {
    const T rdt = T(1) / (T(1) / T(nsplit[0]) * dt);
    const T *rarea = g.rarea;
    const long long gs2 = g.s2;
    if (launch_rect(s, [=] __device__(int i, int j, int z) {
          const int tile = z / npz;
          const T *xk = xfx + (long long)z * d.sx + (long long)(j - d.jsd) * d.nix;
          const T *yk = yfx + (long long)z * d.sy;
          dp1[(long long)z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] =
              (xk[i + 1 - d.is] - xk[i - d.is] + yk[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)] -
               yk[(long long)(j - d.js) * d.nid + (i - d.isd)]) *
              rarea[tile * gs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)] * rdt;
        }, d.is, d.ie, d.js, d.je, npz * nt))
      return -1;
  }
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

template int tracer_2d_plain<double>(Ctx &, bool, double *, double *, double *, double *, double *,
                                     double *, int, int, int, double, int, int, double, double, int *);
template int tracer_2d_plain<float>(Ctx &, bool, float *, float *, float *, float *, float *, float *,
                                    int, int, int, float, int, int, float, float, int *);

// ---- six-tile edge halo ------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(256) k_halo(T *base, long long tile_stride, long long s2,
                                              const long long *dst, const long long *src, int n,
                                              const int *slab_ids, int nslab_dense) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const long long sl = slab_ids ? (long long)slab_ids[blockIdx.y] : (long long)blockIdx.y;
  (void)nslab_dense;
  // dst/src are encoded as tile * tile_marker + cell, tile_marker = 1<<40
  const long long td = dst[e] >> 40, cd = dst[e] & ((1ll << 40) - 1);
  const long long ts = src[e] >> 40, cs = src[e] & ((1ll << 40) - 1);
  base[td * tile_stride + sl * s2 + cd] = base[ts * tile_stride + sl * s2 + cs];
}

template <typename T>
int halo_fill(Ctx &c, T *base, long long tile_stride, const int *slab_ids, int nslab_list,
              int nslab_dense) {
  if (c.ntiles != 6 || c.halo_n == 0) return 0;
  const int nsl = slab_ids ? nslab_list : nslab_dense;
  for (int s0 = 0; s0 < nsl; s0 += 65535) {
    const int ns = std::min(65535, nsl - s0);
    dim3 b(256), g((unsigned)((c.halo_n + 255) / 256), (unsigned)ns);
    hipLaunchKernelGGL(k_halo<T>, g, b, 0, c.stream, base + (slab_ids ? 0 : (long long)s0 * c.d.s2),
                       tile_stride, c.d.s2, c.halo_dst, c.halo_src, c.halo_n,
                       slab_ids ? slab_ids + s0 : nullptr, nslab_dense);
    FV3_HIP(hipGetLastError());
  }
  return 0;
}
template int halo_fill<double>(Ctx &, double *, long long, const int *, int, int);
template int halo_fill<float>(Ctx &, float *, long long, const int *, int, int);

// ---- fv_tp_2d for one 2-D field (parity entry point) ---------------------------
template <typename T>
int fv_tp_2d_plain(Ctx &c, T *q, const T *crx, const T *cry, int hord, T *fx, T *fy,
                   const T *xfx, const T *yfx, const T *ra_x, const T *ra_y, T lim_fac,
                   const T *mfx, const T *mfy, const T *mass, int nord, T damp_c, bool damp) {
  (void)lim_fac;
  const Dims d = c.d;
  const GridDev<T> g = grid_of<T>(c);
  hipStream_t s = c.stream;
  if (hord < 7 || hord > 13) return set_error("hord=%d: only the monotone orders 7..13 are implemented", hord);
  if (damp && (nord < 0 || nord > 2)) return set_error("nord=%d: deln_flux supports nord 0..2", nord);
  const int ord_in = hord == 10 ? 8 : hord, ord_ou = hord;
  const long long nfx2 = (long long)(d.nid + 1) * d.njd, nfy2 = (long long)d.nid * (d.njd + 1);
  const size_t bytes = sizeof(T) * (d.sy + d.sry + d.sx + d.srx + d.s2 + nfx2 + nfy2) + 1024;
  T *t_fy2 = (T *)ctx_scratch(c, bytes);
  if (!t_fy2) return -1;
  T *t_qi = t_fy2 + d.sy, *t_fx2 = t_qi + d.sry, *t_qj = t_fx2 + d.sx, *t_d2 = t_qj + d.srx;
  T *t_dfx = t_d2 + d.s2, *t_dfy = t_dfx + nfx2;
  const LineGeom gx{d.is, d.ie, d.npx, d.edge}, gy{d.js, d.je, d.npy, d.edge};
  const unsigned corners = d.corners;
  const T *area = g.area, *dxa = g.dxa, *dya = g.dya, *rarea = g.rarea;
  auto cc = [&](T *w, int dir) {
    if (d.bounded) return 0;
    return launch_rect(s, [=] __device__(int cidx, int, int) {
          const int blk = cidx / 9, k9 = cidx % 9;
          const int i = (blk == 0 || blk == 3 ? 1 - NG : d.npx) + k9 % 3;
          const int j = (blk == 0 || blk == 1 ? 1 - NG : d.npy) + k9 / 3;
          int si, sj;
          if (corner_src(i, j, d.npx, d.npy, dir, corners, si, sj))
            w[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = w[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
        }, 0, 35, 0, 0, 1);
  };
  if (cc(q, 2)) return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        ColQ<T> qa{q + (i - d.isd), d.jsd, d.nid};
        ColQ<T> da{dya + (i - d.isd), d.jsd, d.nid};
        t_fy2[(long long)(j - d.js) * d.nid + (i - d.isd)] =
            ppm_flux<T>(qa, da, j, cry[(long long)(j - d.js) * d.nid + (i - d.isd)], gy, ord_in);
      }, d.isd, d.ied, d.js, d.je + 1, 1))
    return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        const T fa = yfx[(long long)(j - d.js) * d.nid + (i - d.isd)] * t_fy2[(long long)(j - d.js) * d.nid + (i - d.isd)];
        const T fb = yfx[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)] * t_fy2[(long long)(j + 1 - d.js) * d.nid + (i - d.isd)];
        t_qi[(long long)(j - d.js) * d.nid + (i - d.isd)] =
            (q[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * area[(long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
            ra_y[(long long)(j - d.js) * d.nid + (i - d.isd)];
      }, d.isd, d.ied, d.js, d.je, 1))
    return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        RowQ<T> qa{t_qi + (long long)(j - d.js) * d.nid, d.isd};
        RowQ<T> da{dxa + (long long)(j - d.jsd) * d.nid, d.isd};
        fx[(long long)(j - d.js) * d.nix + (i - d.is)] =
            ppm_flux<T>(qa, da, i, crx[(long long)(j - d.jsd) * d.nix + (i - d.is)], gx, ord_ou);
      }, d.is, d.ie + 1, d.js, d.je, 1))
    return -1;
  if (cc(q, 1)) return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        RowQ<T> qa{q + (long long)(j - d.jsd) * d.nid, d.isd};
        RowQ<T> da{dxa + (long long)(j - d.jsd) * d.nid, d.isd};
        t_fx2[(long long)(j - d.jsd) * d.nix + (i - d.is)] =
            ppm_flux<T>(qa, da, i, crx[(long long)(j - d.jsd) * d.nix + (i - d.is)], gx, ord_in);
      }, d.is, d.ie + 1, d.jsd, d.jed, 1))
    return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        const T *xk = xfx + (long long)(j - d.jsd) * d.nix;
        const T *f2 = t_fx2 + (long long)(j - d.jsd) * d.nix;
        const T fa = xk[i - d.is] * f2[i - d.is];
        const T fb = xk[i + 1 - d.is] * f2[i + 1 - d.is];
        t_qj[(long long)(j - d.jsd) * d.niy + (i - d.is)] =
            (q[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * area[(long long)(j - d.jsd) * d.nid + (i - d.isd)] + fa - fb) /
            ra_x[(long long)(j - d.jsd) * d.niy + (i - d.is)];
      }, d.is, d.ie, d.jsd, d.jed, 1))
    return -1;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        ColQ<T> qa{t_qj + (i - d.is), d.jsd, d.niy};
        ColQ<T> da{dya + (i - d.isd), d.jsd, d.nid};
        fy[(long long)(j - d.js) * d.niy + (i - d.is)] =
            ppm_flux<T>(qa, da, j, cry[(long long)(j - d.js) * d.nid + (i - d.isd)], gy, ord_ou);
      }, d.is, d.ie, d.js, d.je + 1, 1))
    return -1;
  const bool with_mf = mfx != nullptr && mfy != nullptr;
  if (launch_rect(s, [=] __device__(int i, int j, int) {
        if (i <= d.ie + 1 && j <= d.je) {
          const long long o = (long long)(j - d.js) * d.nix + (i - d.is);
          const T wgt = with_mf ? mfx[o] : xfx[(long long)(j - d.jsd) * d.nix + (i - d.is)];
          fx[o] = T(0.5) * (fx[o] + t_fx2[(long long)(j - d.jsd) * d.nix + (i - d.is)]) * wgt;
        }
        if (i <= d.ie && j <= d.je + 1) {
          const long long o = (long long)(j - d.js) * d.niy + (i - d.is);
          const T wgt = with_mf ? mfy[o] : yfx[(long long)(j - d.js) * d.nid + (i - d.isd)];
          fy[o] = T(0.5) * (fy[o] + t_fy2[(long long)(j - d.js) * d.nid + (i - d.isd)]) * wgt;
        }
      }, d.is, d.ie + 1, d.js, d.je + 1, 1))
    return -1;
  const bool do_damp = damp && damp_c > T(1.0e-4) && (with_mf ? mass != nullptr : true);
  if (do_damp) {
    const T dampv = (T)pow((double)damp_c * c.da_min, (double)(nord + 1));
    const bool wmass = with_mf;
    const T *d6u = g.del6_u, *d6v = g.del6_v;
    if (launch_rect(s, [=] __device__(int i, int j, int) {
          const bool in = i >= d.is - 1 - nord && i <= d.ie + 1 + nord && j >= d.js - 1 - nord && j <= d.je + 1 + nord;
          const T qv = q[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
          t_d2[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = in ? (wmass ? qv : dampv * qv) : T(0);
        }, d.isd, d.ied, d.jsd, d.jed, 1))
      return -1;
    auto fx2_pass = [&](int lo, int sign) {
      return launch_rect(s, [=] __device__(int i, int j, int) {
            const T *w = t_d2 + (long long)(j - d.jsd) * d.nid;
            const T dv = sign > 0 ? (w[i - 1 - d.isd] - w[i - d.isd]) : (w[i - d.isd] - w[i - 1 - d.isd]);
            t_dfx[(long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)] = d6v[(long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)] * dv;
          }, d.is - lo, d.ie + lo + 1, d.js - lo, d.je + lo, 1);
    };
    auto fy2_pass = [&](int lo, int sign) {
      return launch_rect(s, [=] __device__(int i, int j, int) {
            const T a = t_d2[(long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)], b = t_d2[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
            t_dfy[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = d6u[(long long)(j - d.jsd) * d.nid + (i - d.isd)] * (sign > 0 ? a - b : b - a);
          }, d.is - lo, d.ie + lo, d.js - lo, d.je + lo + 1, 1);
    };
    if (nord > 0 && cc(t_d2, 1)) return -1;
    if (fx2_pass(nord, 1)) return -1;
    if (nord > 0 && cc(t_d2, 2)) return -1;
    if (fy2_pass(nord, 1)) return -1;
    for (int n = 1; n <= nord; n++) {
      const int ntt = nord - n;
      if (launch_rect(s, [=] __device__(int i, int j, int) {
            const T *fxx = t_dfx + (long long)(j - d.jsd) * (d.nid + 1);
            t_d2[(long long)(j - d.jsd) * d.nid + (i - d.isd)] =
                (fxx[i - d.isd] - fxx[i + 1 - d.isd] + t_dfy[(long long)(j - d.jsd) * d.nid + (i - d.isd)] -
                 t_dfy[(long long)(j + 1 - d.jsd) * d.nid + (i - d.isd)]) *
                rarea[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
          }, d.is - ntt - 1, d.ie + ntt + 1, d.js - ntt - 1, d.je + ntt + 1, 1))
        return -1;
      if (cc(t_d2, 1)) return -1;
      if (fx2_pass(ntt, -1)) return -1;
      if (cc(t_d2, 2)) return -1;
      if (fy2_pass(ntt, -1)) return -1;
    }
    const T damp2 = T(0.5) * dampv;
    if (launch_rect(s, [=] __device__(int i, int j, int) {
          if (i <= d.ie + 1 && j <= d.je) {
            T &f = fx[(long long)(j - d.js) * d.nix + (i - d.is)];
            const T v = t_dfx[(long long)(j - d.jsd) * (d.nid + 1) + (i - d.isd)];
            f = wmass ? f + damp2 * (mass[(long long)(j - d.jsd) * d.nid + (i - 1 - d.isd)] + mass[(long long)(j - d.jsd) * d.nid + (i - d.isd)]) * v
                      : f + v;
          }
          if (i <= d.ie && j <= d.je + 1) {
            T &f = fy[(long long)(j - d.js) * d.niy + (i - d.is)];
            const T v = t_dfy[(long long)(j - d.jsd) * d.nid + (i - d.isd)];
            f = wmass ? f + damp2 * (mass[(long long)(j - 1 - d.jsd) * d.nid + (i - d.isd)] + mass[(long long)(j - d.jsd) * d.nid + (i - d.isd)]) * v
                      : f + v;
          }
        }, d.is, d.ie + 1, d.js, d.je + 1, 1))
      return -1;
  }
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

template int fv_tp_2d_plain<double>(Ctx &, double *, const double *, const double *, int, double *,
                                    double *, const double *, const double *, const double *,
                                    const double *, double, const double *, const double *,
                                    const double *, int, double, bool);

}  // namespace fv3
