// l2e.hip -- the full Lagrangian_to_Eulerian of the vertical remap (upstream
// fv_mapz.F90, reached through fv_dynamics; the reference imports fv_mapz_mod
// at moving_nest/fv_moving_nest.F90:85): beyond the tracers, pt and delp,
// the vertical velocity w (bottom value ws, kord_wz), the layer depth delz
// (remapped as -delz/delp, kord_wz), the D-grid winds u and v (pressure
// interfaces averaged to the cell faces, kord_mt, iv = -1), the Eulerian
// pressure set pe, ps, peln, pk, pkz (moving_nest/fv_moving_nest.F90:
// 2205-2213), the fast-physics hook with the GFDL_interstitial state
// (ccpp/data/CCPP_typedefs.F90:434-504, kmp by the rule at :1714-1718), and
// a total-energy fix-up with compensated FP64 global sums.
//
// Every column goes through the plain column kernel (mapz_column.h: one
// thread per column, upstream evaluation order), so the FP64 form matches a
// CPU FP64 evaluation; fields the kernel does not address in place are
// gathered into [level][column] buffers and scattered back.
#include <algorithm>
#include <cmath>
#include <vector>

#include "../../include/fv3atm_amd.h"
#include "fv3_internal.h"
#include "mapz_column.h"

namespace fv3 {

namespace {

struct Geo {  // extents of one tile's arrays
  int is, ie, js, je, isd, jsd, ni, nj, nid, njd, km;
  long long s2, su, sv, spe, spl, spk, spz;
};

Geo geo_of(const Ctx &c) {
  const Dims &d = c.d;
  Geo g;
  g.is = d.is; g.ie = d.ie; g.js = d.js; g.je = d.je; g.isd = d.isd; g.jsd = d.jsd;
  g.ni = d.ie - d.is + 1; g.nj = d.je - d.js + 1; g.nid = d.nid; g.njd = d.njd; g.km = c.npz;
  g.s2 = d.s2;
  g.su = (long long)d.nid * (d.njd + 1);
  g.sv = (long long)(d.nid + 1) * d.njd;
  g.spe = (long long)(g.ni + 2) * (g.km + 1) * (g.nj + 2);
  g.spl = (long long)g.ni * (g.km + 1) * g.nj;
  g.spk = (long long)g.ni * g.nj * (g.km + 1);
  g.spz = (long long)g.ni * g.nj * g.km;
  return g;
}

// pe (is-1:ie+1, km+1, js-1:je+1) element
__device__ __forceinline__ long long pe_at(const Geo &g, int i, int k, int j) {
  return ((long long)(j - (g.js - 1)) * (g.km + 1) + (k - 1)) * (g.ni + 2) + (i - (g.is - 1));
}

// Lagrangian interfaces and ps on the compute domain and its one-cell ring
__global__ void k_pe_ring(Geo g, int t, double ptop, const double *delp, double *pe, double *ps) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int w = g.ni + 2;
  if (idx >= w * (g.nj + 2)) return;
  const int i = g.is - 1 + idx % w, j = g.js - 1 + idx / w;
  const double *dl = delp + (long long)t * g.km * g.s2 + (long long)(j - g.jsd) * g.nid + (i - g.isd);
  double *p = pe + t * g.spe;
  double v = ptop;
  p[pe_at(g, i, 1, j)] = v;
  for (int k = 1; k <= g.km; k++) {
    v = v + dl[(long long)(k - 1) * g.s2];
    p[pe_at(g, i, k + 1, j)] = v;
  }
  ps[t * g.s2 + (long long)(j - g.jsd) * g.nid + (i - g.isd)] = v;
}

// source/target interfaces [k][col] of the u (y-face), v (x-face) or cell
// columns of tile t; col = (j - js) * ncol_i + (i - is)
__global__ void k_col_pressures(Geo g, int t, int kind, double ptop, const double *ak, const double *bk,
                                const double *pe, const double *ps, double *pe1, double *pe2, long long nc) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= nc) return;
  const int wi = kind == 2 ? g.ni + 1 : g.ni;
  const int i = g.is + (int)(col % wi), j = g.js + (int)(col / wi);
  const double *p = pe + t * g.spe;
  const double *s = ps + t * g.s2;
  auto PS = [&](int ii, int jj) { return s[(long long)(jj - g.jsd) * g.nid + (ii - g.isd)]; };
  for (int k = 1; k <= g.km + 1; k++) {
    double a, b;
    if (kind == 1) {  // u: cells j-1, j
      const double bkh = 0.5 * bk[k - 1];
      a = 0.5 * (p[pe_at(g, i, k, j - 1)] + p[pe_at(g, i, k, j)]);
      b = ak[k - 1] + bkh * (PS(i, j - 1) + PS(i, j));
    } else if (kind == 2) {  // v: cells i-1, i
      const double bkh = 0.5 * bk[k - 1];
      a = 0.5 * (p[pe_at(g, i - 1, k, j)] + p[pe_at(g, i, k, j)]);
      b = ak[k - 1] + bkh * (PS(i - 1, j) + PS(i, j));
    } else {  // cell column
      a = p[pe_at(g, i, k, j)];
      b = k == 1 ? ptop : (k == g.km + 1 ? PS(i, j) : ak[k - 1] + bk[k - 1] * PS(i, j));
    }
    pe1[(long long)(k - 1) * nc + col] = a;
    pe2[(long long)(k - 1) * nc + col] = b;
  }
}

// gather (dir 0) or scatter (dir 1) a field's compute-domain columns:
// field element (i, j, k) at k*lev + (j - j0)*ld + (i - i0); buffer [k][col]
template <typename T>
__global__ void k_gather(T *field, T *buf, int dir, int km, int wi, long long nc, long long lev, int ld,
                         int di, int dj) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nc * km) return;
  const long long col = idx % nc;
  const int k = (int)(idx / nc);
  const int ii = (int)(col % wi) + di, jj = (int)(col / wi) + dj;
  T *f = field + (long long)k * lev + (long long)jj * ld + ii;
  if (dir == 0)
    buf[idx] = *f;
  else
    *f = buf[idx];
}

// delz <-> -delz/delp (before: dp is the Lagrangian delp; after: dp2 from pe2)
template <typename T>
__global__ void k_delz(Geo g, int t, int dir, T *dz, const double *delp, const double *pe2) {
  const long long nc = (long long)g.ni * g.nj;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nc * g.km) return;
  const long long col = idx % nc;
  const int k = (int)(idx / nc) + 1;
  const int i = g.is + (int)(col % g.ni), j = g.js + (int)(col / g.ni);
  T *p = dz + t * g.spz + idx;
  if (dir == 0)
    *p = -*p / T(delp[((long long)t * g.km + (k - 1)) * g.s2 + (long long)(j - g.jsd) * g.nid + (i - g.isd)]);
  else
    *p = -*p * T(pe2[(long long)k * nc + col] - pe2[(long long)(k - 1) * nc + col]);
}

// column total energy sum_k (cp pt pkz + (ua^2 + va^2)/2) dp; dp from delp
template <typename T>
__device__ double column_te(const Geo &g, int t, int i, int j, double cp, const T *u, const T *v, const T *pt,
                            const double *pkz, const double *delp) {
  double te = 0.0;
  const long long cell = (long long)(j - g.jsd) * g.nid + (i - g.isd);
  const long long pc = (long long)(j - g.js) * g.ni + (i - g.is);
  for (int k = 1; k <= g.km; k++) {
    const T *uk = u + ((long long)t * g.km + (k - 1)) * g.su;
    const T *vk = v + ((long long)t * g.km + (k - 1)) * g.sv;
    const double ua = 0.5 * ((double)uk[(long long)(j - g.jsd) * g.nid + (i - g.isd)] +
                             (double)uk[(long long)(j + 1 - g.jsd) * g.nid + (i - g.isd)]);
    const double va = 0.5 * ((double)vk[(long long)(j - g.jsd) * (g.nid + 1) + (i - g.isd)] +
                             (double)vk[(long long)(j - g.jsd) * (g.nid + 1) + (i + 1 - g.isd)]);
    const double ptk = (double)pt[((long long)t * g.km + (k - 1)) * g.s2 + cell];
    const double pz = pkz[t * g.spz + (long long)(k - 1) * g.ni * g.nj + pc];
    const double d = delp[((long long)t * g.km + (k - 1)) * g.s2 + cell];
    te = te + (cp * ptk * pz + 0.5 * (ua * ua + va * va)) * d;
  }
  return te;
}

template <typename T>
__global__ void k_te(Geo g, int t, double cp, const T *u, const T *v, const T *pt, const double *pkz,
                     const double *delp, double *te) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ni * g.nj) return;
  const int i = g.is + idx % g.ni, j = g.js + idx / g.ni;
  te[(long long)t * g.ni * g.nj + idx] = column_te<T>(g, t, i, j, cp, u, v, pt, pkz, delp);
}

// Eulerian pe (compute domain), peln, pk, pkz
template <typename T>
__global__ void k_pressures(Geo g, int t, double ptop, const double *ak, const double *bk, double akap,
                            double rrg, int hydrostatic, const double *ps, const double *delp, const T *pt,
                            const T *delz, double *pe, double *peln, double *pk, double *pkz) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ni * g.nj) return;
  const int i = g.is + idx % g.ni, j = g.js + idx / g.ni;
  const long long cell = (long long)(j - g.jsd) * g.nid + (i - g.isd);
  const double psv = ps[t * g.s2 + cell];
  const long long nc = (long long)g.ni * g.nj;
  double *pl = peln + t * g.spl + (long long)(j - g.js) * (g.km + 1) * g.ni + (i - g.is);
  double *pkc = pk + t * g.spk + idx;
  for (int k = 1; k <= g.km + 1; k++) {
    const double p = k == 1 ? ptop : (k == g.km + 1 ? psv : ak[k - 1] + bk[k - 1] * psv);
    pe[t * g.spe + pe_at(g, i, k, j)] = p;
    const double lp = log(p);
    pl[(long long)(k - 1) * g.ni] = lp;
    pkc[(long long)(k - 1) * nc] = exp(akap * lp);
  }
  for (int k = 1; k <= g.km; k++) {
    double *o = pkz + t * g.spz + (long long)(k - 1) * nc + idx;
    if (hydrostatic) {
      *o = (pkc[(long long)k * nc] - pkc[(long long)(k - 1) * nc]) /
           (akap * (pl[(long long)k * g.ni] - pl[(long long)(k - 1) * g.ni]));
    } else {
      const double dz = (double)delz[t * g.spz + (long long)(k - 1) * nc + idx];
      const double ptk = (double)pt[((long long)t * g.km + (k - 1)) * g.s2 + cell];
      // p = rho R pt p^akap, so p^akap = (rho R pt)^(akap / (1 - akap))
      const double k1k = akap / (1.0 - akap);
      *o = exp(k1k * log(rrg * delp[((long long)t * g.km + (k - 1)) * g.s2 + cell] / dz * ptk));
    }
  }
}

// Neumaier sums of x[i] * w[i] (w may be NULL) over n values, deterministic:
// block partials of fixed strided order, then one block combines them.
struct KS {
  double s, c;
};
__device__ __forceinline__ void kadd(KS &k, double x) {
  const double t = k.s + x;
  k.c += fabs(k.s) >= fabs(x) ? (k.s - t) + x : (x - t) + k.s;
  k.s = t;
}
__device__ __forceinline__ void kmerge(KS &a, const KS &b) {
  kadd(a, b.s);
  a.c += b.c;
}

constexpr int KS_BLOCK = 256, KS_GRID = 256;

__global__ void __launch_bounds__(KS_BLOCK) k_ksum_part(const double *x, const double *w, long long n, KS *part) {
  __shared__ KS sh[KS_BLOCK];
  KS k{0.0, 0.0};
  for (long long i = (long long)blockIdx.x * KS_BLOCK + threadIdx.x; i < n; i += (long long)KS_BLOCK * KS_GRID)
    kadd(k, w ? x[i] * w[i] : x[i]);
  sh[threadIdx.x] = k;
  __syncthreads();
  for (int s = KS_BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) kmerge(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_ksum_final(const KS *part, int np, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    KS k{0.0, 0.0};
    for (int i = 0; i < np; i++) kmerge(k, part[i]);
    *out = k.s + k.c;
  }
}

// per-cell terms of the energy fix-up: te1 * area, cp (ps - ptop) area, te0 * area
template <typename T>
__global__ void k_energy_terms(Geo g, int t, double cp, double ptop, const T *u, const T *v, const T *pt,
                               const double *pkz, const double *delp, const double *ps, const double *area,
                               const double *te0, double *e0, double *e1, double *m) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.ni * g.nj) return;
  const int i = g.is + idx % g.ni, j = g.js + idx / g.ni;
  const long long cell = (long long)(j - g.jsd) * g.nid + (i - g.isd);
  const double ar = area[t * g.s2 + cell];
  const long long o = (long long)t * g.ni * g.nj + idx;
  e0[o] = te0[o] * ar;
  e1[o] = column_te<T>(g, t, i, j, cp, u, v, pt, pkz, delp) * ar;
  m[o] = cp * (ps[t * g.s2 + cell] - ptop) * ar;
}

__global__ void k_dtmp(const double *sums, double consv, double *dtmp) {
  if (threadIdx.x == 0 && blockIdx.x == 0) dtmp[0] = consv * (sums[0] - sums[1]) / sums[2];
}

template <typename T>
__global__ void k_add_dtmp(Geo g, int nt, T *pt, const double *pkz, const double *dtmp) {
  const long long nc = (long long)g.ni * g.nj;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nc * g.km * nt) return;
  const long long col = idx % nc;
  const int k = (int)((idx / nc) % g.km) + 1, t = (int)(idx / (nc * g.km));
  const int i = g.is + (int)(col % g.ni), j = g.js + (int)(col / g.ni);
  T *p = pt + ((long long)t * g.km + (k - 1)) * g.s2 + (long long)(j - g.jsd) * g.nid + (i - g.isd);
  *p = *p + T(dtmp[0] / pkz[t * g.spz + (long long)(k - 1) * nc + col]);
}

dim3 grid1(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// the l2e arena: buffers that stay alive across the column-kernel calls
// (which use the context's scratch arena for their own work)
void *l2e_arena(Ctx &c, size_t bytes) {
  if (bytes <= c.l2e_bytes) return c.l2e_buf;
  if (c.graph_safe && c.l2e_buf) {
    set_error("graph-safe mode: the Lagrangian_to_Eulerian arena cannot grow");
    return nullptr;
  }
  if (c.l2e_buf) {
    (void)hipStreamSynchronize(c.stream);
    (void)hipFree(c.l2e_buf);
    c.l2e_buf = nullptr;
    c.l2e_bytes = 0;
  }
  if (hipMalloc(&c.l2e_buf, bytes) != hipSuccess) {
    set_error("Lagrangian_to_Eulerian arena of %zu bytes failed", bytes);
    c.l2e_buf = nullptr;
    return nullptr;
  }
  c.l2e_bytes = bytes;
  return c.l2e_buf;
}

}  // namespace

// pfull / kmp rule of GFDL_interstitial (ccpp/data/CCPP_typedefs.F90:1706-1718):
// the first level whose reference full-level pressure exceeds 10 hPa
int l2e_kmp(int npz, double p_ref, const double *ak, const double *bk) {
  int kmp = npz;
  for (int k = 1; k <= npz; k++) {
    const double ph1 = ak[k - 1] + bk[k - 1] * p_ref, ph2 = ak[k] + bk[k] * p_ref;
    const double pfull = (ph2 - ph1) / std::log(ph2 / ph1);
    kmp = k;
    if (pfull > 10.0e2) break;
  }
  return kmp;
}

template <typename T>
int l2e_full(Ctx &c, const fv3_l2e_args &a) {
  const Geo g = geo_of(c);
  const int km = c.npz, nt = c.ntiles;
  hipStream_t s = c.stream;
  if (km < 4) return set_error("Lagrangian_to_Eulerian needs km >= 4 (got %d)", km);
  auto kord_ok = [](int k) { return k > 7 && (k <= 10 || k > 16); };
  if (!kord_ok(a.kord_mt) || (!a.hydrostatic && !kord_ok(a.kord_wz)))
    return set_error("kord_mt=%d kord_wz=%d: implemented orders are 8, 9, 10 and > 16", a.kord_mt, a.kord_wz);
  if (!a.delp || !a.u || !a.v || !a.pt || !a.pe || !a.peln || !a.pk || !a.pkz || !a.ps || !a.ak || !a.bk)
    return set_error("Lagrangian_to_Eulerian: NULL array argument");
  if (!a.hydrostatic && (!a.w || !a.delz || !a.ws))
    return set_error("Lagrangian_to_Eulerian: the non-hydrostatic remap needs w, delz and ws");
  if (a.nq > 0 && (!a.q || !a.kord_tr)) return set_error("Lagrangian_to_Eulerian: NULL q or kord_tr");
  if (a.consv > 0.0 && !a.te0_2d) return set_error("the energy fix-up (consv > 0) needs te0_2d");
  const bool energy = a.te0_2d != nullptr;
  T *u = (T *)a.u, *v = (T *)a.v, *w = (T *)a.w, *dz = (T *)a.delz, *pt = (T *)a.pt, *q = (T *)a.q;
  // arena: per-kind column pressures (the largest column set: (ni+1)*(nj+1)),
  // a gathered field, energy terms, sums
  const long long ncmax = (long long)(g.ni + 1) * (g.nj + 1);
  const long long ncell = (long long)g.ni * g.nj * nt;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_pe1 = take(sizeof(double) * (km + 1) * ncmax), o_pe2 = take(sizeof(double) * (km + 1) * ncmax);
  const size_t o_buf = take(sizeof(T) * km * ncmax), o_ws = take(sizeof(T) * ncmax);
  const size_t o_e = take(sizeof(double) * 3 * ncell), o_part = take(sizeof(KS) * KS_GRID * 3);
  const size_t o_sum = take(sizeof(double) * 4);
  char *base = (char *)l2e_arena(c, off);
  if (!base) return 1;
  double *pe1 = (double *)(base + o_pe1), *pe2 = (double *)(base + o_pe2);
  T *buf = (T *)(base + o_buf), *wsb = (T *)(base + o_ws);
  double *e0 = (double *)(base + o_e), *e1 = e0 + ncell, *em = e1 + ncell;
  KS *part = (KS *)(base + o_part);
  double *sums = (double *)(base + o_sum);

  for (int t = 0; t < nt; t++) {
    hipLaunchKernelGGL(k_pe_ring, grid1((long long)(g.ni + 2) * (g.nj + 2)), dim3(256), 0, s, g, t, a.ptop, a.delp,
                       a.pe, a.ps);
    if (energy)
      hipLaunchKernelGGL(k_te<T>, grid1((long long)g.ni * g.nj), dim3(256), 0, s, g, t, a.cp, u, v, pt, a.pkz,
                         a.delp, a.te0_2d);
  }
  FV3_HIP(hipGetLastError());

  // one field of tile t over the columns of a kind through the column kernel
  auto remap_kind = [&](int t, int kind, T *field, long long lev, int ld, int di, int dj, int iv, int kord,
                        const T *qs) -> int {
    const int wi = kind == 2 ? g.ni + 1 : g.ni, wj = kind == 1 ? g.nj + 1 : g.nj;
    const long long nc = (long long)wi * wj;
    hipLaunchKernelGGL(k_col_pressures, grid1(nc), dim3(256), 0, s, g, t, kind, a.ptop, a.ak, a.bk, a.pe, a.ps,
                       pe1, pe2, nc);
    hipLaunchKernelGGL(k_gather<T>, grid1(nc * km), dim3(256), 0, s, field, buf, 0, km, wi, nc, lev, ld, di, dj);
    FV3_HIP(hipGetLastError());
    RemapArgs<T> r{};
    r.mode = 1;
    r.km = km;
    r.kn = km;
    r.ncol = nc;
    r.pe1_in = pe1;
    r.pe2_in = pe2;
    r.q1_in = buf;
    r.q2_out = buf;
    r.qs_in = qs;
    r.q1_stride = nc;
    r.q2_stride = nc;
    r.iv = iv;
    r.kord1 = kord;
    r.variant = 0;
    if (remap_launch<T>(c, r)) return 1;
    hipLaunchKernelGGL(k_gather<T>, grid1(nc * km), dim3(256), 0, s, field, buf, 1, km, wi, nc, lev, ld, di, dj);
    FV3_HIP(hipGetLastError());
    return 0;
  };
  for (int t = 0; t < nt; t++) {
    // u (isd:ied, jsd:jed+1): columns is:ie x js:je+1
    if (remap_kind(t, 1, u + (long long)t * km * g.su, g.su, g.nid, g.is - g.isd, g.js - g.jsd, -1, a.kord_mt,
                   nullptr))
      return 1;
    // v (isd:ied+1, jsd:jed): columns is:ie+1 x js:je
    if (remap_kind(t, 2, v + (long long)t * km * g.sv, g.sv, g.nid + 1, g.is - g.isd, g.js - g.jsd, -1,
                   a.kord_mt, nullptr))
      return 1;
    if (!a.hydrostatic) {
      // w: cell columns with the surface value ws as the bottom boundary
      FV3_HIP(hipMemcpyAsync(wsb, (const T *)a.ws + (long long)t * g.ni * g.nj, sizeof(T) * g.ni * g.nj,
                             hipMemcpyDeviceToDevice, s));
      if (remap_kind(t, 0, w + (long long)t * km * g.s2, g.s2, g.nid, g.is - g.isd, g.js - g.jsd, -2, a.kord_wz,
                     wsb))
        return 1;
      // delz as -delz/delp over the cell columns (delz is (is:ie, js:je, km))
      hipLaunchKernelGGL(k_delz<T>, grid1((long long)g.ni * g.nj * km), dim3(256), 0, s, g, t, 0, dz, a.delp, pe2);
      if (remap_kind(t, 0, dz + t * g.spz, (long long)g.ni * g.nj, g.ni, 0, 0, 1, a.kord_wz, nullptr)) return 1;
      hipLaunchKernelGGL(k_delz<T>, grid1((long long)g.ni * g.nj * km), dim3(256), 0, s, g, t, 1, dz, a.delp, pe2);
      FV3_HIP(hipGetLastError());
    }
  }
  // tracers, pt, delp <- dp2
  {
    std::vector<int> kt(std::max(1, a.nq), 0);
    for (int iq = 0; iq < a.nq; iq++) kt[iq] = a.kord_tr[iq];
    if (remap_tracers_plain<T>(c, km, a.nq, a.ak, a.bk, a.ptop, a.delp, q, pt, kt.data(), a.kord_tm, a.t_min,
                               a.fill))
      return 1;
  }
  const double rrg = -a.rdgas / a.grav;
  for (int t = 0; t < nt; t++)
    hipLaunchKernelGGL(k_pressures<T>, grid1((long long)g.ni * g.nj), dim3(256), 0, s, g, t, a.ptop, a.ak, a.bk,
                       a.akap, rrg, a.hydrostatic, a.ps, a.delp, pt, dz, a.pe, a.peln, a.pk, a.pkz);
  FV3_HIP(hipGetLastError());
  // fast physics: the caller's hook sees the Eulerian state
  if (a.fast_physics) {
    fv3_l2e_state st{};
    st.npz = km;
    st.nq = a.nq;
    st.ntiles = nt;
    std::vector<double> akh(km + 1), bkh(km + 1);
    FV3_HIP(hipMemcpyAsync(akh.data(), a.ak, sizeof(double) * (km + 1), hipMemcpyDeviceToHost, s));
    FV3_HIP(hipMemcpyAsync(bkh.data(), a.bk, sizeof(double) * (km + 1), hipMemcpyDeviceToHost, s));
    FV3_HIP(hipStreamSynchronize(s));
    st.kmp = l2e_kmp(km, a.p_ref > 0.0 ? a.p_ref : 1.0e5, akh.data(), bkh.data());
    st.last_step = a.last_step;
    st.mdt = a.mdt;
    st.delp = a.delp;
    st.pt = a.pt;
    st.q = a.q;
    st.delz = a.delz;
    st.w = a.w;
    st.pkz = a.pkz;
    st.pk = a.pk;
    st.peln = a.peln;
    st.pe = a.pe;
    st.ps = a.ps;
    st.te0_2d = a.te0_2d;
    st.stream = (void *)s;
    if (a.fast_physics(a.fast_physics_user, &st)) return set_error("the fast-physics hook reported failure");
  }
  // total-energy fix-up: the domain's energy loss back into pt
  if (energy && a.consv > 0.0) {
    for (int t = 0; t < nt; t++)
      hipLaunchKernelGGL(k_energy_terms<T>, grid1((long long)g.ni * g.nj), dim3(256), 0, s, g, t, a.cp, a.ptop, u,
                         v, pt, a.pkz, a.delp, a.ps, c.grid64.area, a.te0_2d, e0, e1, em);
    for (int m = 0; m < 3; m++) {
      hipLaunchKernelGGL(k_ksum_part, dim3(KS_GRID), dim3(KS_BLOCK), 0, s, e0 + m * ncell, (const double *)nullptr,
                         ncell, part + m * KS_GRID);
      hipLaunchKernelGGL(k_ksum_final, dim3(1), dim3(1), 0, s, part + m * KS_GRID, KS_GRID, sums + m);
    }
    hipLaunchKernelGGL(k_dtmp, dim3(1), dim3(1), 0, s, sums, a.consv, sums + 3);
    hipLaunchKernelGGL(k_add_dtmp<T>, grid1((long long)g.ni * g.nj * km * nt), dim3(256), 0, s, g, nt, pt, a.pkz,
                       sums + 3);
    FV3_HIP(hipGetLastError());
    if (a.dtmp_out) FV3_HIP(hipMemcpyAsync(a.dtmp_out, sums + 3, sizeof(double), hipMemcpyDeviceToHost, s));
  }
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

template int l2e_full<double>(Ctx &, const fv3_l2e_args &);
template int l2e_full<float>(Ctx &, const fv3_l2e_args &);

}  // namespace fv3

using namespace fv3;

extern "C" {

int fv3_lagrangian_to_eulerian_dev(fv3_ctx *h, int form, const fv3_l2e_args *a, void *stream) {
  if (!h || !a) return set_error("fv3_lagrangian_to_eulerian_dev: NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.grid_set) return set_error("grid metrics not uploaded (fv3_ctx_set_grid)");
  if (c.graph_safe) return set_error("Lagrangian_to_Eulerian synchronises the host; turn graph-safe mode off");
  if (a->km != c.npz) return set_error("km=%d differs from the context npz=%d", a->km, c.npz);
  if (a->nq < 0 || a->nq > REMAP_MAX_NQ) return set_error("nq=%d out of 0..%d", a->nq, REMAP_MAX_NQ);
  hipStream_t saved = c.stream;
  if (stream) c.stream = (hipStream_t)stream;
  int rc;
  if (form == FV3_FORM_PLAIN64)
    rc = l2e_full<double>(c, *a);
  else if (form == FV3_FORM_FUSED32)
    rc = l2e_full<float>(c, *a);
  else
    rc = set_error("form %d: Lagrangian_to_Eulerian runs FV3_FORM_PLAIN64 (double fields) or "
                   "FV3_FORM_FUSED32 (float fields)", form);
  c.stream = saved;
  return rc;
}

int fv3_kmp(int npz, double p_ref, const double *ak, const double *bk, int *kmp) {
  if (!ak || !bk || !kmp || npz < 1) return set_error("fv3_kmp: bad argument");
  *kmp = l2e_kmp(npz, p_ref, ak, bk);
  return 0;
}

}  // extern "C"
