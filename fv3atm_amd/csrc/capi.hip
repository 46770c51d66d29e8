// capi.hip -- the extern "C" boundary of libfv3atm_amd.so: context, metric
// upload, host-array entry points mirroring the upstream Fortran interfaces
// (staged through device buffers), and the device-resident step.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/fv3atm_amd.h"
#include "fused.h"
#include "fv3_internal.h"
#include "mapz_fast.h"
#include "mapz_column.h"

namespace fv3 {

static thread_local char g_err[1024] = "";

int set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return 1;
}

Dims make_dims(int is, int ie, int js, int je, int isd, int ied, int jsd, int jed, int npx,
               int npy, unsigned corners, bool bounded, int grid_type) {
  Dims d{};
  d.is = is;
  d.ie = ie;
  d.js = js;
  d.je = je;
  d.isd = isd;
  d.ied = ied;
  d.jsd = jsd;
  d.jed = jed;
  d.npx = npx;
  d.npy = npy;
  d.nid = ied - isd + 1;
  d.njd = jed - jsd + 1;
  d.nix = ie - is + 2;
  d.niy = ie - is + 1;
  d.nj = je - js + 1;
  d.njy = je - js + 2;
  d.s2 = (long long)d.nid * d.njd;
  d.sx = (long long)d.nix * d.njd;
  d.sy = (long long)d.nid * d.njy;
  d.smx = (long long)d.nix * d.nj;
  d.smy = (long long)d.niy * d.njy;
  d.srx = (long long)d.niy * d.njd;
  d.sry = (long long)d.nid * d.nj;
  d.sdp = (long long)d.niy * d.nj;
  d.corners = corners;
  d.bounded = bounded;
  d.edge = !bounded && grid_type < 3;
  return d;
}

// A graph-safe context may be referenced by captured graphs: an arena that
// exists is never freed or moved while graph-safe steps are on (a call that
// needs it bigger fails instead); one that does not exist yet may be made.
static bool arena_frozen(const Ctx &c, const void *p, size_t have, size_t want, const char *what) {
  if (c.graph_safe && p && want > have) {
    set_error("the context is in graph-safe mode and its %s arena (%zu bytes) cannot grow to %zu "
              "bytes: captured graphs may hold its address (turn graph-safe mode off first, or "
              "size the context with a larger eager call before capturing)", what, have, want);
    return true;
  }
  return false;
}

void *ctx_scratch(Ctx &c, size_t bytes) {
  if (arena_frozen(c, c.scratch, c.scratch_bytes, bytes, "scratch")) return nullptr;
  if (bytes > c.scratch_bytes) {
    if (c.scratch) {
      (void)hipStreamSynchronize(c.stream);
      (void)hipFree(c.scratch);
      c.scratch = nullptr;
      c.scratch_bytes = 0;
    }
    hipError_t e = hipMalloc(&c.scratch, bytes);
    if (e != hipSuccess) {
      set_error("scratch allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      c.scratch = nullptr;
      return nullptr;
    }
    c.scratch_bytes = bytes;
  }
  return c.scratch;
}

void *ctx_stage(Ctx &c, size_t bytes) {
  if (arena_frozen(c, c.stage, c.stage_bytes, bytes, "staging")) return nullptr;
  if (bytes > c.stage_bytes) {
    if (c.stage) {
      (void)hipStreamSynchronize(c.stream);
      (void)hipFree(c.stage);
      c.stage = nullptr;
      c.stage_bytes = 0;
    }
    hipError_t e = hipMalloc(&c.stage, bytes);
    if (e != hipSuccess) {
      set_error("staging allocation of %zu bytes failed: %s", bytes, hipGetErrorString(e));
      c.stage = nullptr;
      return nullptr;
    }
    c.stage_bytes = bytes;
  }
  return c.stage;
}

template <>
const GridDev<double> &grid_of<double>(const Ctx &c) {
  return c.grid64;
}
template <>
const GridDev<float> &grid_of<float>(const Ctx &c) {
  return c.grid32;
}

// FMS cubic-mosaic contacts: the source cell of every edge-halo cell.
static void build_halo_table(int n, std::vector<long long> &dst, std::vector<long long> &src) {
  const int isd = 1 - NG, nid = n + 2 * NG;
  auto cell = [&](int i, int j) { return (long long)(j - isd) * nid + (i - isd); };
  auto nxt = [](int t, int s) { return ((t - 1 + s) % 6 + 6) % 6 + 1; };
  auto add = [&](int t, int i, int j, int ts, int si, int sj) {
    dst.push_back(((long long)(t - 1) << 40) | cell(i, j));
    src.push_back(((long long)(ts - 1) << 40) | cell(si, sj));
  };
  for (int t = 1; t <= 6; t++) {
    const bool odd = t % 2 == 1;
    for (int e = 1; e <= NG; e++)
      for (int a = 1; a <= n; a++) {
        if (odd) {
          add(t, n + e, a, nxt(t, 1), e, a);
          add(t, a, n + e, nxt(t, 2), e, n + 1 - a);
          add(t, 1 - e, a, nxt(t, -2), n + 1 - a, n + 1 - e);
          add(t, a, 1 - e, nxt(t, -1), a, n + 1 - e);
        } else {
          add(t, n + e, a, nxt(t, 2), n + 1 - a, e);
          add(t, a, n + e, nxt(t, 1), a, e);
          add(t, 1 - e, a, nxt(t, -1), n + 1 - e, a);
          add(t, a, 1 - e, nxt(t, -2), n + 1 - e, n + 1 - a);
        }
      }
  }
}

// default context for the stateless row entry points (map1_ppm & co.)
static Ctx *row_ctx() {
  static std::once_flag once;
  static Ctx *c = nullptr;
  std::call_once(once, [] {
    c = new Ctx();
    (void)hipStreamCreate(&c->stream);
  });
  return c;
}

}  // namespace fv3

using namespace fv3;

extern "C" {

const char *fv3_last_error(void) { return fv3::g_err; }

int fv3_ctx_create(const fv3_bounds *bd, int npx, int npy, int npz, int ntiles, fv3_ctx **out) {
  if (!bd || !out) return set_error("fv3_ctx_create: NULL argument");
  *out = nullptr;
  if (ntiles != 1 && ntiles != 6) return set_error("ntiles must be 1 or 6 (got %d)", ntiles);
  if (npz < 1) return set_error("npz must be >= 1 (got %d)", npz);
  if (bd->is - bd->isd != NG || bd->ied - bd->ie != NG || bd->js - bd->jsd != NG || bd->jed - bd->je != NG)
    return set_error("bounds must carry a %d-cell halo", NG);
  if (bd->ie - bd->is + 1 < 8 || bd->je - bd->js + 1 < 8)
    return set_error("compute domain must be at least 8x8");
  if (ntiles == 6 && (bd->is != 1 || bd->js != 1 || bd->ie != npx - 1 || bd->je != npy - 1 || npx != npy))
    return set_error("a six-tile context needs whole square tiles (is=js=1, ie=je=npx-1)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error("no HIP device visible");
  auto *h = new fv3_ctx();
  Ctx &c = h->c;
  c.npz = npz;
  c.ntiles = ntiles;
  (void)hipGetDevice(&c.dev);
  if (noalloc_setup(c.dev) || remap_rows_setup()) {
    delete h;
    return 1;
  }
  c.d = make_dims(bd->is, bd->ie, bd->js, bd->je, bd->isd, bd->ied, bd->jsd, bd->jed, npx, npy, 15u,
                  false, 0);
  if (hipStreamCreate(&c.stream) != hipSuccess) {
    delete h;
    return set_error("hipStreamCreate failed");
  }
  for (auto &a : c.aux)
    if (hipStreamCreateWithFlags(&a, hipStreamNonBlocking) != hipSuccess) {
      fv3_ctx_destroy(h);
      return set_error("hipStreamCreateWithFlags failed");
    }
  if (hipEventCreateWithFlags(&c.fork_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.join_ev[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.join_ev[1], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.join_ev[2], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.prep_ev, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c.chunk_ev, hipEventDisableTiming) != hipSuccess) {
    fv3_ctx_destroy(h);
    return set_error("hipEventCreateWithFlags failed");
  }
  const long long s2 = c.d.s2, sdx = (long long)c.d.nid * (c.d.njd + 1), sdy = (long long)(c.d.nid + 1) * c.d.njd;
  const long long per_tile = s2 * 4 + sdx * 2 + sdy * 2 + s2 * 9;
  if (hipMalloc(&c.g64, sizeof(double) * per_tile * ntiles) != hipSuccess ||
      hipMalloc(&c.g32, sizeof(float) * per_tile * ntiles) != hipSuccess) {
    fv3_ctx_destroy(h);
    return set_error("metric allocation failed");
  }
  auto lay = [&](auto *base, auto &gd) {
    using T = std::remove_pointer_t<decltype(base)>;
    T *p = base;
    gd.area = p;
    p += s2 * ntiles;
    gd.rarea = p;
    p += s2 * ntiles;
    gd.dxa = p;
    p += s2 * ntiles;
    gd.dya = p;
    p += s2 * ntiles;
    gd.dx = p;
    p += sdx * ntiles;
    gd.del6_u = p;
    p += sdx * ntiles;
    gd.dy = p;
    p += sdy * ntiles;
    gd.del6_v = p;
    p += sdy * ntiles;
    gd.sin_sg = p;
    gd.s2 = s2;
    gd.sdx = sdx;
    gd.sdy = sdy;
  };
  lay(c.g64, c.grid64);
  lay(c.g32, c.grid32);
  if (ntiles == 6) {
    std::vector<long long> dst, src;
    build_halo_table(npx - 1, dst, src);
    c.halo_n = (int)dst.size();
    if (hipMalloc(&c.halo_dst, sizeof(long long) * dst.size()) != hipSuccess ||
        hipMalloc(&c.halo_src, sizeof(long long) * src.size()) != hipSuccess) {
      fv3_ctx_destroy(h);
      return set_error("halo table allocation failed");
    }
    (void)hipMemcpy(c.halo_dst, dst.data(), sizeof(long long) * dst.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(c.halo_src, src.data(), sizeof(long long) * src.size(), hipMemcpyHostToDevice);
  }
  *out = h;
  return 0;
}

int fv3_ctx_destroy(fv3_ctx *h) {
  if (!h) return 0;
  Ctx &c = h->c;
  if (c.stream) (void)hipStreamSynchronize(c.stream);
  (void)hipFree(c.g64);
  (void)hipFree(c.g32);
  (void)hipFree(c.gfac32);
  (void)hipFree(c.halo_dst);
  (void)hipFree(c.halo_src);
  if (c.halo_adj) (void)hipFree(c.halo_adj);
  (void)hipFree(c.scratch);
  (void)hipFree(c.stage);
  (void)hipFree(c.step_flag);
  for (auto *t : c.halo_tab) (void)hipFree(t);
  (void)hipFree(c.l2e_buf);
  for (auto &e : c.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &a : c.aux)
    if (a) {
      (void)hipStreamSynchronize(a);
      (void)hipStreamDestroy(a);
    }
  if (c.fork_ev) (void)hipEventDestroy(c.fork_ev);
  if (c.prep_ev) (void)hipEventDestroy(c.prep_ev);
  if (c.chunk_ev) (void)hipEventDestroy(c.chunk_ev);
  for (auto &e : c.join_ev)
    if (e) (void)hipEventDestroy(e);
  if (c.stream) (void)hipStreamDestroy(c.stream);
  delete h;
  return 0;
}

int fv3_ctx_set_grid(fv3_ctx *h, int tile, const fv3_grid *g) {
  if (!h || !g) return set_error("fv3_ctx_set_grid: NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (tile < 1 || tile > c.ntiles) return set_error("tile %d out of 1..%d", tile, c.ntiles);
  const long long s2 = c.d.s2, sdx = c.grid64.sdx, sdy = c.grid64.sdy;
  const int t = tile - 1;
  struct F {
    const double *src;
    double *d64;
    float *d32;
    long long n;
  } f[] = {
      {g->area, (double *)c.grid64.area + t * s2, (float *)c.grid32.area + t * s2, s2},
      {g->rarea, (double *)c.grid64.rarea + t * s2, (float *)c.grid32.rarea + t * s2, s2},
      {g->dxa, (double *)c.grid64.dxa + t * s2, (float *)c.grid32.dxa + t * s2, s2},
      {g->dya, (double *)c.grid64.dya + t * s2, (float *)c.grid32.dya + t * s2, s2},
      {g->dx, (double *)c.grid64.dx + t * sdx, (float *)c.grid32.dx + t * sdx, sdx},
      {g->del6_u, (double *)c.grid64.del6_u + t * sdx, (float *)c.grid32.del6_u + t * sdx, sdx},
      {g->dy, (double *)c.grid64.dy + t * sdy, (float *)c.grid32.dy + t * sdy, sdy},
      {g->del6_v, (double *)c.grid64.del6_v + t * sdy, (float *)c.grid32.del6_v + t * sdy, sdy},
      {g->sin_sg, (double *)c.grid64.sin_sg + t * 9 * s2, (float *)c.grid32.sin_sg + t * 9 * s2, 9 * s2},
  };
  std::vector<float> tmp;
  for (auto &x : f) {
    if (!x.src) return set_error("fv3_ctx_set_grid: a metric array is NULL");
    FV3_HIP(hipMemcpy(x.d64, x.src, sizeof(double) * x.n, hipMemcpyHostToDevice));
    tmp.resize(x.n);
    for (long long k = 0; k < x.n; k++) tmp[k] = (float)x.src[k];
    FV3_HIP(hipMemcpy(x.d32, tmp.data(), sizeof(float) * x.n, hipMemcpyHostToDevice));
  }
  {  // FP32 combined upwind factors of this tile for the fused FP32 form
    const int nid = c.d.nid, njd = c.d.njd;
    const long long nx = (long long)(nid + 1) * njd, ny = (long long)nid * (njd + 1);
    std::vector<float> f(2 * nx + 2 * ny + s2, 0.0f);
    for (int j = 0; j < njd; j++)
      for (int i = 1; i < nid; i++) {  // x-face i between cells i-1 and i (0-based from isd)
        const long long fo = (long long)j * (nid + 1) + i;
        const double dyv = g->dy[fo];
        f[fo] = (float)(g->dxa[(long long)j * nid + i - 1] * dyv * g->sin_sg[2 * s2 + (long long)j * nid + i - 1]);
        f[nx + fo] = (float)(g->dxa[(long long)j * nid + i] * dyv * g->sin_sg[0 * s2 + (long long)j * nid + i]);
      }
    for (int j = 1; j < njd; j++)
      for (int i = 0; i < nid; i++) {  // y-face j between cells j-1 and j
        const long long fo = (long long)j * nid + i;
        const double dxv = g->dx[fo];
        f[2 * nx + fo] = (float)(g->dya[(long long)(j - 1) * nid + i] * dxv * g->sin_sg[3 * s2 + (long long)(j - 1) * nid + i]);
        f[2 * nx + ny + fo] = (float)(g->dya[(long long)j * nid + i] * dxv * g->sin_sg[1 * s2 + (long long)j * nid + i]);
      }
    // the cube-edge face value of xppm between cells e-1 and e (e = is at the
    // left edge, ie + 1 at the right one) as a weighted sum of cells e-2..e+1:
    // 1/2 [((2 d(e-1) + d(e-2)) q(e-1) - d(e-1) q(e-2)) / (d(e-2) + d(e-1))
    //    + ((2 d(e) + d(e+1)) q(e) - d(e) q(e+1)) / (d(e) + d(e+1))], d = dxa
    if (g->bounded_domain == 0 && g->grid_type < 3)
      for (int j = 0; j < njd; j++)
        for (int e : {c.d.is - c.d.isd, c.d.ie + 1 - c.d.isd}) {
          if (e < 2 || e + 1 >= nid) continue;
          const double *dr = g->dxa + (long long)j * nid;
          const double da = dr[e - 2], db = dr[e - 1], dc = dr[e], de = dr[e + 1];
          float *w = f.data() + 2 * nx + 2 * ny + (long long)j * nid;
          w[e - 2] = (float)(-0.5 * db / (da + db));
          w[e - 1] = (float)(0.5 * (2.0 * db + da) / (da + db));
          w[e] = (float)(0.5 * (2.0 * dc + de) / (dc + de));
          w[e + 1] = (float)(-0.5 * dc / (dc + de));
        }
    if (!c.gfac32) FV3_HIP(hipMalloc(&c.gfac32, sizeof(float) * f.size() * c.ntiles));
    FV3_HIP(hipMemcpy(c.gfac32 + (long long)t * f.size(), f.data(), sizeof(float) * f.size(),
                      hipMemcpyHostToDevice));
  }
  if (tile == 1 || g->da_min < c.da_min || c.da_min == 0.0) c.da_min = g->da_min;
  unsigned corners = (g->sw_corner ? 1u : 0u) | (g->se_corner ? 2u : 0u) | (g->ne_corner ? 4u : 0u) |
                     (g->nw_corner ? 8u : 0u);
  c.d = make_dims(c.d.is, c.d.ie, c.d.js, c.d.je, c.d.isd, c.d.ied, c.d.jsd, c.d.jed, c.d.npx, c.d.npy,
                  corners, g->bounded_domain != 0, g->grid_type);
  c.grid_set = true;
  return 0;
}

// ---- host-array entry points --------------------------------------------------
struct HostArr {
  void *host;
  size_t bytes;
  bool in, out;
};

static int stage_run(Ctx &c, std::vector<HostArr> arrs, std::vector<void *> &dev) {
  size_t total = 0;
  for (auto &a : arrs) total += (a.bytes + 255) & ~(size_t)255;
  char *base = (char *)ctx_stage(c, total + 256);
  if (!base) return -1;
  size_t off = 0;
  dev.clear();
  for (auto &a : arrs) {
    dev.push_back(a.host ? base + off : nullptr);
    if (a.host && a.in) FV3_HIP(hipMemcpyAsync(base + off, a.host, a.bytes, hipMemcpyHostToDevice, c.stream));
    off += (a.bytes + 255) & ~(size_t)255;
  }
  return 0;
}

static int stage_back(Ctx &c, const std::vector<HostArr> &arrs, const std::vector<void *> &dev) {
  for (size_t k = 0; k < arrs.size(); k++)
    if (arrs[k].host && arrs[k].out)
      FV3_HIP(hipMemcpyAsync(arrs[k].host, dev[k], arrs[k].bytes, hipMemcpyDeviceToHost, c.stream));
  FV3_HIP(hipStreamSynchronize(c.stream));
  return 0;
}

static int check_ctx(fv3_ctx *h, int npx, int npy, int npz) {
  if (!h) return set_error("NULL context");
  if (!h->c.grid_set) return set_error("grid metrics not uploaded (fv3_ctx_set_grid)");
  if (npx != h->c.d.npx || npy != h->c.d.npy) return set_error("npx/npy differ from the context");
  if (npz > 0 && npz != h->c.npz) return set_error("npz=%d differs from the context (%d)", npz, h->c.npz);
  return 0;
}

int fv3_fv_tp_2d(fv3_ctx *h, double *q, const double *crx, const double *cry, int npx, int npy,
                 int hord, double *fx, double *fy, const double *xfx, const double *yfx,
                 const double *ra_x, const double *ra_y, double lim_fac, const double *mfx,
                 const double *mfy, const double *mass, const int *nord, const double *damp_c) {
  if (check_ctx(h, npx, npy, 0)) return 1;
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (c.ntiles != 1) return set_error("fv3_fv_tp_2d takes a single-tile context");
  if (!q || !crx || !cry || !fx || !fy || !xfx || !yfx || !ra_x || !ra_y)
    return set_error("fv3_fv_tp_2d: NULL required argument");
  if ((mfx == nullptr) != (mfy == nullptr)) return set_error("mfx and mfy must be given together");
  const Dims &d = c.d;
  std::vector<HostArr> arrs = {
      {q, sizeof(double) * d.s2, true, true},          {(void *)crx, sizeof(double) * d.sx, true, false},
      {(void *)cry, sizeof(double) * d.sy, true, false}, {fx, sizeof(double) * d.smx, false, true},
      {fy, sizeof(double) * d.smy, false, true},       {(void *)xfx, sizeof(double) * d.sx, true, false},
      {(void *)yfx, sizeof(double) * d.sy, true, false}, {(void *)ra_x, sizeof(double) * d.srx, true, false},
      {(void *)ra_y, sizeof(double) * d.sry, true, false}, {(void *)mfx, sizeof(double) * d.smx, true, false},
      {(void *)mfy, sizeof(double) * d.smy, true, false}, {(void *)mass, sizeof(double) * d.s2, true, false},
  };
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  const bool damp = nord != nullptr && damp_c != nullptr;
  if (fv_tp_2d_plain<double>(c, (double *)dv[0], (double *)dv[1], (double *)dv[2], hord, (double *)dv[3],
                             (double *)dv[4], (double *)dv[5], (double *)dv[6], (double *)dv[7],
                             (double *)dv[8], lim_fac, (double *)dv[9], (double *)dv[10],
                             (double *)dv[11], damp ? *nord : 0, damp ? *damp_c : 0.0, damp))
    return 1;
  return stage_back(c, arrs, dv);
}

}  // extern "C"

// ---- the host-array transport through the fused kernels ------------------------
// After a fused transport (which reads the winds and dp1 without changing
// them), the in-place semantics of tracer_2d: the Courant numbers and mass
// fluxes scaled by 1/nsplit(k), dp1 advanced to the thickness at the start
// of the last sub-cycle (the oracle's running sum, in its order).
template <typename T>
__global__ void k_scale_levels(T *a, long long per_level, int npz, int nt, const int *nsplit) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= per_level * npz * nt) return;
  const int k = (int)((idx / per_level) % npz);
  const int ns = nsplit[k];
  if (ns > 1) {
    const double frac = 1.0 / (double)ns;
    a[idx] = a[idx] * (T)frac;
  }
}

template <typename T>
__global__ void k_dp1_advance(T *dp1, const T *mfx, const T *mfy, const double *rarea, Dims d, int npz, int nt,
                              const int *nsplit, long long grs2) {
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ni * nj * npz * nt) return;
  const int i = d.is + (int)(idx % ni), j = d.js + (int)((idx / ni) % nj);
  const int k = (int)((idx / ((long long)ni * nj)) % npz), t = (int)(idx / ((long long)ni * nj * npz));
  const int ns = nsplit[k];
  if (ns < 2) return;
  const long long z = (long long)t * npz + k;
  const T *mx = mfx + z * d.smx + (long long)(j - d.js) * d.nix;
  const T *my = mfy + z * d.smy;
  const double ra = rarea[t * grs2 + (long long)(j - d.jsd) * d.nid + (i - d.isd)];
  T *p = dp1 + z * d.s2 + (long long)(j - d.jsd) * d.nid + (i - d.isd);
  T v = *p;
  for (int it = 1; it < ns; it++)
    v = v + (T)(((mx[i - d.is] - mx[i + 1 - d.is]) +
                 (my[(long long)(j - d.js) * d.niy + (i - d.is)] - my[(long long)(j + 1 - d.js) * d.niy + (i - d.is)])) *
                (T)ra);
  *p = v;
}

// fv_tp_2d runs copy_corners on its q argument (tp_core.F90 fv_tp_2d: dir 2
// before the y sweep, dir 1 before the x sweep), so a level the caller's q
// went through in place leaves the transport with its cube-corner halo
// blocks holding the dir-1 substitution of the adjacent halo strips: every
// level of tracer_2d, and the levels with nsplit = 1 of tracer_2d_1L (whose
// sub-cycled levels run on a copy, fv_tracer2d.F90 tracer_2d_1L qn2).  The
// strips themselves are not written by the transport.
template <typename T>
__global__ void k_corner_blocks(T *q, Dims d, int npz, int nq, int nt, const int *nsplit, int per_level) {
  constexpr int nb = NG * NG;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= 4LL * nb * npz * nq * nt) return;
  const int cell = (int)(idx % (4 * nb)), corner = cell / nb, a = cell % NG, b = (cell / NG) % NG;
  const long long lev = idx / (4 * nb);
  const int k = (int)(lev % npz);
  if (per_level && nsplit[k] != 1) return;
  const int i = (corner == 1 || corner == 2) ? d.npx + a : 1 - NG + a;
  const int j = corner >= 2 ? d.npy + b : 1 - NG + b;
  int si, sj;
  if (!fv3::corner_src(i, j, d.npx, d.npy, 1, d.corners, si, sj)) return;
  if (i < d.isd || i > d.ied || j < d.jsd || j > d.jed || si < d.isd || si > d.ied || sj < d.jsd || sj > d.jed)
    return;
  T *p = q + lev * d.s2;
  p[(long long)(j - d.jsd) * d.nid + (i - d.isd)] = p[(long long)(sj - d.jsd) * d.nid + (si - d.isd)];
}

template <typename A, typename B>
__global__ void k_convert(const A *a, B *b, long long n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx < n) b[idx] = (B)a[idx];
}

static dim3 grid_of_n(long long n) { return dim3((unsigned)((n + 255) / 256)); }

// bytes the fused step keeps at the front of the staging arena (Courant
// maxima, fractions, sub-cycle counts, flag)
static size_t step_head_bytes(int npz) { return 4 * (((size_t)8 * npz + 255) & ~(size_t)255) + 4096; }

// tracer_2d / tracer_2d_1L on host arrays of T through the fused transport
template <typename T>
static int fused_transport_host(Ctx &c, bool per_level, int q_split, T *q, T *dp1, T *mfx, T *mfy, T *cx,
                                T *cy, int nq, int hord, int *nsplit_out) {
  const Dims &d = c.d;
  const int npz = c.npz, nt = c.ntiles;
  hipStream_t s = c.stream;
  const long long nqv = d.s2 * npz * nq * nt, ndp = d.s2 * npz * nt;
  const long long ncx = d.sx * npz * nt, ncy = d.sy * npz * nt, nmx = d.smx * npz * nt, nmy = d.smy * npz * nt;
  size_t off = step_head_bytes(npz);
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  const size_t oq = take(sizeof(T) * nqv), oqo = take(sizeof(T) * nqv), odp = take(sizeof(T) * ndp);
  const size_t odd = take(sizeof(double) * ndp), ocx = take(sizeof(T) * ncx), ocy = take(sizeof(T) * ncy);
  const size_t omx = take(sizeof(T) * nmx), omy = take(sizeof(T) * nmy), ons = take(sizeof(int) * npz);
  char *base = (char *)ctx_stage(c, off);
  if (!base) return 1;
  T *dq = (T *)(base + oq), *dqo = (T *)(base + oqo), *ddp = (T *)(base + odp);
  double *ddd = (double *)(base + odd);
  T *dcx = (T *)(base + ocx), *dcy = (T *)(base + ocy), *dmx = (T *)(base + omx), *dmy = (T *)(base + omy);
  int *dns = (int *)(base + ons);
  FV3_HIP(hipMemcpyAsync(dq, q, sizeof(T) * nqv, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(ddp, dp1, sizeof(T) * ndp, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dcx, cx, sizeof(T) * ncx, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dcy, cy, sizeof(T) * ncy, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dmx, mfx, sizeof(T) * nmx, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dmy, mfy, sizeof(T) * nmy, hipMemcpyHostToDevice, s));
  // the step writes the interior of q_out only: its halo keeps q's
  FV3_HIP(hipMemcpyAsync(dqo, dq, sizeof(T) * nqv, hipMemcpyDeviceToDevice, s));
  const double *dp1d = (const double *)ddp;
  if (sizeof(T) != sizeof(double)) {
    hipLaunchKernelGGL((k_convert<T, double>), grid_of_n(ndp), dim3(256), 0, s, ddp, ddd, ndp);
    FV3_HIP(hipGetLastError());
    dp1d = ddd;
  }
  const bool saved_pl = c.step_per_level;
  const int saved_qs = c.step_q_split;
  c.step_per_level = per_level;
  c.step_q_split = per_level ? 0 : q_split;
  c.transport_only = true;
  const int rc = fused_step(c, sizeof(T) == 4, nq, dq, dqo, nullptr, dp1d, nullptr, dcx, dcy, dmx, dmy, nullptr,
                            nullptr, 0.0, hord, 9, 10, 184.0, 0);
  c.transport_only = false;
  c.step_per_level = saved_pl;
  c.step_q_split = saved_qs;
  if (rc) return rc;
  if ((int)c.last_nsplit.size() != npz) return set_error("sub-cycle counts of the transport missing");
  FV3_HIP(hipMemcpyAsync(dns, c.last_nsplit.data(), sizeof(int) * npz, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_scale_levels<T>, grid_of_n(ncx), dim3(256), 0, s, dcx, d.sx, npz, nt, dns);
  hipLaunchKernelGGL(k_scale_levels<T>, grid_of_n(ncy), dim3(256), 0, s, dcy, d.sy, npz, nt, dns);
  hipLaunchKernelGGL(k_scale_levels<T>, grid_of_n(nmx), dim3(256), 0, s, dmx, d.smx, npz, nt, dns);
  hipLaunchKernelGGL(k_scale_levels<T>, grid_of_n(nmy), dim3(256), 0, s, dmy, d.smy, npz, nt, dns);
  const long long ncell = (long long)(d.ie - d.is + 1) * (d.je - d.js + 1) * npz * nt;
  hipLaunchKernelGGL(k_dp1_advance<T>, grid_of_n(ncell), dim3(256), 0, s, ddp, dmx, dmy, c.grid64.rarea, d, npz,
                     nt, dns, c.grid64.s2);
  if (!d.bounded)
    hipLaunchKernelGGL(k_corner_blocks<T>, grid_of_n(4LL * NG * NG * npz * nq * nt), dim3(256), 0, s, dqo, d, npz,
                       nq, nt, dns, per_level ? 1 : 0);
  FV3_HIP(hipGetLastError());
  FV3_HIP(hipMemcpyAsync(q, dqo, sizeof(T) * nqv, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(dp1, ddp, sizeof(T) * ndp, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(cx, dcx, sizeof(T) * ncx, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(cy, dcy, sizeof(T) * ncy, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(mfx, dmx, sizeof(T) * nmx, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(mfy, dmy, sizeof(T) * nmy, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  if (nsplit_out)
    for (int k = 0; k < npz; k++) nsplit_out[k] = c.last_nsplit[k];
  return 0;
}

extern "C" {

static int tracer_common(fv3_ctx *h, bool per_level, double *q, double *dp1, double *mfx,
                         double *mfy, double *cx, double *cy, int npx, int npy, int npz, int nq,
                         int hord, int q_split, double dt, int id_divg, int nord_tr, double trdm,
                         double lim_fac, int *nsplit_out) {
  if (check_ctx(h, npx, npy, npz)) return 1;
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!q || !dp1 || !mfx || !mfy || !cx || !cy) return set_error("tracer_2d: NULL array argument");
  if (nq < 1) return set_error("nq must be >= 1 (got %d)", nq);
  if (q_split < 0) return set_error("q_split must be >= 0 (got %d)", q_split);
  // the fused FP64 transport (bit for bit the plain form's results) unless
  // the call asks for what only the plain form does: deln_flux damping on
  // the first sub-cycle, dp1 overwritten with the divergence (id_divg), or
  // tracer_2d on several faces, whose q halo the reference leaves holding
  // the exchange before the last sub-cycle
  const bool damp = !per_level && trdm > 1.0e-4;
  const bool divg = !per_level && id_divg > 0;
  const bool faces = !per_level && c.ntiles > 1;
  if (!damp && !divg && !faces && hord >= 7 && hord <= 13)
    return fused_transport_host<double>(c, per_level, q_split, q, dp1, mfx, mfy, cx, cy, nq, hord, nsplit_out);
  const Dims &d = c.d;
  const long long nt = c.ntiles;
  std::vector<HostArr> arrs = {
      {q, sizeof(double) * d.s2 * npz * nq * nt, true, true}, {dp1, sizeof(double) * d.s2 * npz * nt, true, true},
      {mfx, sizeof(double) * d.smx * npz * nt, true, true},  {mfy, sizeof(double) * d.smy * npz * nt, true, true},
      {cx, sizeof(double) * d.sx * npz * nt, true, true},    {cy, sizeof(double) * d.sy * npz * nt, true, true},
  };
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  if (tracer_2d_plain<double>(c, per_level, (double *)dv[0], (double *)dv[1], (double *)dv[2], (double *)dv[3],
                              (double *)dv[4], (double *)dv[5], nq, hord, q_split, dt, id_divg, nord_tr,
                              trdm, lim_fac, nsplit_out))
    return 1;
  return stage_back(c, arrs, dv);
}

int fv3_tracer_2d_1l(fv3_ctx *h, double *q, double *dp1, double *mfx, double *mfy, double *cx,
                     double *cy, int npx, int npy, int npz, int nq, int hord, int q_split,
                     double dt, int id_divg, int nord_tr, double trdm, double lim_fac,
                     int *nsplit_out) {
  return tracer_common(h, true, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split, dt,
                       id_divg, nord_tr, trdm, lim_fac, nsplit_out);
}

int fv3_tracer_2d(fv3_ctx *h, double *q, double *dp1, double *mfx, double *mfy, double *cx,
                  double *cy, int npx, int npy, int npz, int nq, int hord, int q_split, double dt,
                  int id_divg, int nord_tr, double trdm, double lim_fac, int *nsplit_out) {
  return tracer_common(h, false, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split, dt,
                       id_divg, nord_tr, trdm, lim_fac, nsplit_out);
}

static int tracer_common_r4(fv3_ctx *h, bool per_level, float *q, float *dp1, float *mfx, float *mfy,
                            float *cx, float *cy, int npx, int npy, int npz, int nq, int hord, int q_split,
                            int id_divg, int nord_tr, double trdm, int *nsplit_out);

int fv3_tracer_2d_1l_r4(fv3_ctx *h, float *q, float *dp1, float *mfx, float *mfy, float *cx, float *cy,
                        int npx, int npy, int npz, int nq, int hord, int q_split, double dt, int id_divg,
                        int nord_tr, double trdm, double lim_fac, int *nsplit_out) {
  (void)dt;
  (void)lim_fac;
  return tracer_common_r4(h, true, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split, id_divg,
                          nord_tr, trdm, nsplit_out);
}

int fv3_tracer_2d_r4(fv3_ctx *h, float *q, float *dp1, float *mfx, float *mfy, float *cx, float *cy, int npx,
                     int npy, int npz, int nq, int hord, int q_split, double dt, int id_divg, int nord_tr,
                     double trdm, double lim_fac, int *nsplit_out) {
  (void)dt;
  (void)lim_fac;
  return tracer_common_r4(h, false, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split, id_divg,
                          nord_tr, trdm, nsplit_out);
}

int fv3_lagrangian_to_eulerian_tracers_r4(fv3_ctx *h, int km, int nq, const double *ak, const double *bk,
                                          double ptop, float *delp, float *q, float *pt, const int *kord_tr,
                                          int kord_tm, double t_min, int fill) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (km != c.npz) return set_error("km=%d differs from the context npz=%d", km, c.npz);
  if (!ak || !bk || !delp || nq < 1 || !q || !kord_tr) return set_error("remap: NULL array argument or nq < 1");
  if (fill) return set_error("the 32-bit remap has no fillz (fill must be 0)");
  for (int iq = 1; iq < nq; iq++)
    if (kord_tr[iq] != kord_tr[0]) return set_error("the 32-bit remap takes one kord_tr for every tracer");
  const Dims &d = c.d;
  const int nt = c.ntiles;
  hipStream_t s = c.stream;
  const long long nqv = d.s2 * km * nq * nt, ndp = d.s2 * km * nt;
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  const size_t oq = take(sizeof(float) * nqv), oqo = take(sizeof(float) * nqv), odp = take(sizeof(float) * ndp);
  const size_t od1 = take(sizeof(double) * ndp), od2 = take(sizeof(double) * ndp);
  const size_t opt = take(sizeof(float) * ndp), oab = take(sizeof(double) * 2 * (km + 1));
  char *base = (char *)ctx_stage(c, off);
  if (!base) return 1;
  float *dq = (float *)(base + oq), *dqo = (float *)(base + oqo), *ddp = (float *)(base + odp);
  double *dp1d = (double *)(base + od1), *dp2d = (double *)(base + od2), *dab = (double *)(base + oab);
  float *dpt = pt ? (float *)(base + opt) : nullptr;
  FV3_HIP(hipMemcpyAsync(dq, q, sizeof(float) * nqv, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dqo, dq, sizeof(float) * nqv, hipMemcpyDeviceToDevice, s));
  FV3_HIP(hipMemcpyAsync(ddp, delp, sizeof(float) * ndp, hipMemcpyHostToDevice, s));
  if (pt) FV3_HIP(hipMemcpyAsync(dpt, pt, sizeof(float) * ndp, hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dab, ak, sizeof(double) * (km + 1), hipMemcpyHostToDevice, s));
  FV3_HIP(hipMemcpyAsync(dab + km + 1, bk, sizeof(double) * (km + 1), hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL((k_convert<float, double>), grid_of_n(ndp), dim3(256), 0, s, ddp, dp1d, ndp);
  FV3_HIP(hipGetLastError());
  if (remap_only_f32(c, nq, dq, dqo, dpt, dp1d, dp2d, dab, dab + km + 1, ptop, kord_tr[0], kord_tm, t_min)) return 1;
  hipLaunchKernelGGL((k_convert<double, float>), grid_of_n(ndp), dim3(256), 0, s, dp2d, ddp, ndp);
  FV3_HIP(hipGetLastError());
  FV3_HIP(hipMemcpyAsync(q, dqo, sizeof(float) * nqv, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipMemcpyAsync(delp, ddp, sizeof(float) * ndp, hipMemcpyDeviceToHost, s));
  if (pt) FV3_HIP(hipMemcpyAsync(pt, dpt, sizeof(float) * ndp, hipMemcpyDeviceToHost, s));
  FV3_HIP(hipStreamSynchronize(s));
  return 0;
}

static int tracer_common_r4(fv3_ctx *h, bool per_level, float *q, float *dp1, float *mfx, float *mfy,
                            float *cx, float *cy, int npx, int npy, int npz, int nq, int hord, int q_split,
                            int id_divg, int nord_tr, double trdm, int *nsplit_out) {
  if (check_ctx(h, npx, npy, npz)) return 1;
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!q || !dp1 || !mfx || !mfy || !cx || !cy) return set_error("tracer_2d: NULL array argument");
  if (nq < 1) return set_error("nq must be >= 1 (got %d)", nq);
  if (q_split < 0) return set_error("q_split must be >= 0 (got %d)", q_split);
  if (hord < 7 || hord > 13) return set_error("hord=%d: only the monotone orders 7..13 are implemented", hord);
  if ((!per_level && trdm > 1.0e-4) || (!per_level && id_divg > 0) || nord_tr < 0)
    return set_error("the 32-bit form has no deln_flux damping or id_divg (trdm=%g id_divg=%d)", trdm, id_divg);
  return fused_transport_host<float>(c, per_level, q_split, q, dp1, mfx, mfy, cx, cy, nq, hord, nsplit_out);
}

// ---- row entry points (stateless) ---------------------------------------------
static bool kord_ok(int kord) { return kord > 7 && (kord <= 10 || kord > 16); }

static int row_map(int variant, int km, const double *pe1, const double *q1, const double *qs,
                   int kn, const double *pe2, double *q2, const double *dp2, int i1, int i2,
                   int j, int ibeg, int iend, int jbeg, int jend, int iv, int kord, bool use_qmin,
                   double qmin) {
  if (!pe1 || !q1 || !pe2 || !q2) return set_error("map: NULL array argument");
  if (km < 4 || kn < 1) return set_error("map: km >= 4 and kn >= 1 required");
  if (!kord_ok(kord)) return set_error("kord=%d: implemented orders are 8, 9, 10 and > 16", kord);
  if (i1 > i2 || i1 < ibeg || i2 > iend || j < jbeg || j > jend) return set_error("map: bounds out of range");
  if (iv == -2 && !qs) return set_error("map: iv=-2 needs qs");
  Ctx &c = *row_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const long long im = i2 - i1 + 1, ni = iend - ibeg + 1, plane = ni * (jend - jbeg + 1);
  // gather the row's columns contiguously
  std::vector<double> q1r(im * km), pe1h(pe1, pe1 + im * (km + 1)), pe2h(pe2, pe2 + im * (kn + 1));
  for (int k = 0; k < km; k++)
    for (long long i = 0; i < im; i++) q1r[k * im + i] = q1[k * plane + (long long)(j - jbeg) * ni + (i1 - ibeg) + i];
  std::vector<HostArr> arrs = {
      {pe1h.data(), sizeof(double) * im * (km + 1), true, false},
      {pe2h.data(), sizeof(double) * im * (kn + 1), true, false},
      {q1r.data(), sizeof(double) * im * km, true, false},
      {(void *)qs, sizeof(double) * im, true, false},
      {(void *)dp2, sizeof(double) * im * kn, true, false},
  };
  std::vector<double> q2r(im * kn);
  arrs.push_back({q2r.data(), sizeof(double) * im * kn, false, true});
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  RemapArgs<double> a{};
  a.mode = 1;
  a.km = km;
  a.kn = kn;
  a.ncol = im;
  a.pe1_in = (const double *)dv[0];
  a.pe2_in = (const double *)dv[1];
  a.q1_in = (const double *)dv[2];
  a.qs_in = (const double *)dv[3];
  a.dp2_in = (const double *)dv[4];
  a.q2_out = (double *)dv[5];
  a.q1_stride = im;
  a.q2_stride = im;
  a.iv = iv;
  a.kord1 = kord;
  a.variant = variant;
  a.use_qmin = use_qmin;
  a.qmin = qmin;
  if (remap_launch<double>(c, a)) return 1;
  if (stage_back(c, arrs, dv)) return 1;
  for (int k = 0; k < kn; k++)
    for (long long i = 0; i < im; i++) q2[k * plane + (long long)(j - jbeg) * ni + (i1 - ibeg) + i] = q2r[k * im + i];
  return 0;
}

int fv3_map1_ppm(int km, const double *pe1, const double *q1, const double *qs, int kn,
                 const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                 int jbeg, int jend, int iv, int kord) {
  return row_map(0, km, pe1, q1, qs, kn, pe2, q2, nullptr, i1, i2, j, ibeg, iend, jbeg, jend, iv,
                 kord, false, 0.0);
}

int fv3_map_scalar(int km, const double *pe1, const double *q1, const double *qs, int kn,
                   const double *pe2, double *q2, int i1, int i2, int j, int ibeg, int iend,
                   int jbeg, int jend, int iv, int kord, double q_min) {
  return row_map(1, km, pe1, q1, qs, kn, pe2, q2, nullptr, i1, i2, j, ibeg, iend, jbeg, jend, iv,
                 kord, true, q_min);
}

int fv3_mapn_tracer(int nq, int km, const double *pe1, const double *pe2, double *q1,
                    const double *dp2, const int *kord, int j, int i1, int i2, int isd, int ied,
                    int jsd, int jed, double q_min, int fill) {
  if (!pe1 || !pe2 || !q1 || !dp2 || !kord) return set_error("mapn_tracer: NULL array argument");
  if (nq < 1 || nq > REMAP_MAX_NQ) return set_error("mapn_tracer: nq=%d out of 1..%d", nq, REMAP_MAX_NQ);
  if (km < 4) return set_error("mapn_tracer: km >= 4 required");
  for (int iq = 0; iq < nq; iq++)
    if (!kord_ok(kord[iq])) return set_error("kord(%d)=%d: implemented orders are 8, 9, 10 and > 16", iq + 1, kord[iq]);
  if (i1 > i2 || i1 < isd || i2 > ied || j < jsd || j > jed) return set_error("mapn_tracer: bounds out of range");
  Ctx &c = *row_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const long long im = i2 - i1 + 1, ni = ied - isd + 1, plane = ni * (jed - jsd + 1);
  std::vector<double> qr(im * km * nq), pe1h(pe1, pe1 + im * (km + 1)), pe2h(pe2, pe2 + im * (km + 1));
  for (int iq = 0; iq < nq; iq++)
    for (int k = 0; k < km; k++)
      for (long long i = 0; i < im; i++)
        qr[((long long)iq * km + k) * im + i] = q1[((long long)iq * km + k) * plane + (long long)(j - jsd) * ni + (i1 - isd) + i];
  std::vector<double> dp2h(dp2, dp2 + im * km);
  std::vector<HostArr> arrs = {
      {pe1h.data(), sizeof(double) * im * (km + 1), true, false},
      {pe2h.data(), sizeof(double) * im * (km + 1), true, false},
      {qr.data(), sizeof(double) * im * km * nq, true, true},
      {dp2h.data(), sizeof(double) * im * km, true, false},
  };
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  // one tracer at a time through the mapn_tracer arithmetic (mode 1, variant 3)
  for (int iq = 0; iq < nq; iq++) {
    RemapArgs<double> a{};
    a.mode = 1;
    a.km = km;
    a.kn = km;
    a.ncol = im;
    a.pe1_in = (const double *)dv[0];
    a.pe2_in = (const double *)dv[1];
    a.q1_in = (const double *)dv[2] + (long long)iq * km * im;
    a.q2_out = (double *)dv[2] + (long long)iq * km * im;
    a.dp2_in = (const double *)dv[3];
    a.q1_stride = im;
    a.q2_stride = im;
    a.iv = 0;
    a.kord1 = kord[iq];
    a.variant = 3;
    a.use_qmin = true;
    a.qmin = q_min;
    a.fill = fill;
    if (remap_launch<double>(c, a)) return 1;
  }
  if (stage_back(c, arrs, dv)) return 1;
  for (int iq = 0; iq < nq; iq++)
    for (int k = 0; k < km; k++)
      for (long long i = 0; i < im; i++)
        q1[((long long)iq * km + k) * plane + (long long)(j - jsd) * ni + (i1 - isd) + i] = qr[((long long)iq * km + k) * im + i];
  return 0;
}

int fv3_cs_profile(const double *qs, double *a4, const double *delp, int km, int i1, int i2,
                   int iv, int kord) {
  if (!a4 || !delp) return set_error("cs_profile: NULL array argument");
  if (km < 4 || i1 > i2) return set_error("cs_profile: km >= 4 and i1 <= i2 required");
  if (!(kord_ok(kord) || (kord >= 1 && kord < 9))) return set_error("kord=%d not implemented", kord);
  if (iv == -2 && !qs) return set_error("cs_profile: iv=-2 needs qs");
  Ctx &c = *row_ctx();
  std::lock_guard<std::mutex> lk(c.mu);
  const long long im = i2 - i1 + 1;
  std::vector<double> a1(im * km);
  for (int k = 0; k < km; k++)
    for (long long i = 0; i < im; i++) a1[k * im + i] = a4[((long long)k * im + i) * 4 + 0];
  std::vector<double> dph(delp, delp + im * km);
  std::vector<double> out(4 * im * km);
  std::vector<HostArr> arrs = {
      {a1.data(), sizeof(double) * im * km, true, false},
      {dph.data(), sizeof(double) * im * km, true, false},
      {(void *)qs, sizeof(double) * im, true, false},
      {out.data(), sizeof(double) * 4 * im * km, false, true},
  };
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  if (cs_profile_rows(c, (const double *)dv[0], (const double *)dv[1], (const double *)dv[2],
                      (double *)dv[3], (int)im, km, iv, kord))
    return 1;
  if (stage_back(c, arrs, dv)) return 1;
  for (int k = 0; k < km; k++)
    for (long long i = 0; i < im; i++)
      for (int m = 0; m < 4; m++) a4[((long long)k * im + i) * 4 + m] = out[((long long)m * km + k) * im + i];
  return 0;
}

int fv3_lagrangian_to_eulerian_tracers(fv3_ctx *h, int km, int nq, const double *ak,
                                       const double *bk, double ptop, double *delp, double *q,
                                       double *pt, const int *kord_tr, int kord_tm, double t_min,
                                       int fill) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (km != c.npz) return set_error("km=%d differs from the context npz=%d", km, c.npz);
  if (!ak || !bk || !delp || (nq > 0 && (!q || !kord_tr))) return set_error("remap: NULL array argument");
  const Dims &d = c.d;
  const long long nt = c.ntiles;
  std::vector<HostArr> arrs = {
      {(void *)ak, sizeof(double) * (km + 1), true, false},
      {(void *)bk, sizeof(double) * (km + 1), true, false},
      {delp, sizeof(double) * d.s2 * km * nt, true, true},
      {q, sizeof(double) * d.s2 * km * nq * nt, true, true},
      {pt, sizeof(double) * d.s2 * km * nt, true, true},
  };
  std::vector<void *> dv;
  if (stage_run(c, arrs, dv)) return 1;
  if (remap_tracers_plain<double>(c, km, nq, (const double *)dv[0], (const double *)dv[1], ptop,
                                  (double *)dv[2], (double *)dv[3], (double *)dv[4], kord_tr,
                                  kord_tm, t_min, fill))
    return 1;
  return stage_back(c, arrs, dv);
}

// ---- device-resident step -------------------------------------------------------
int fv3_step_dev(fv3_ctx *h, int form, int nq, const void *q_in, void *q_out, void *pt,
                 const double *dp1, double *delp_out, const void *cx, const void *cy,
                 const void *mfx, const void *mfy, const double *ak, const double *bk,
                 double ptop, int hord, int kord_tr, int kord_tm, double t_min, int nsplit_cap,
                 void *stream) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.grid_set) return set_error("grid metrics not uploaded (fv3_ctx_set_grid)");
  if (!q_in || !q_out || !dp1 || !delp_out || !cx || !cy || !mfx || !mfy || !ak || !bk)
    return set_error("fv3_step_dev: NULL argument");
  if (q_in == q_out) return set_error("fv3_step_dev: q_in and q_out must be distinct buffers");
  if (nq < 1 || nq > REMAP_MAX_NQ) return set_error("nq=%d out of 1..%d", nq, REMAP_MAX_NQ);
  hipStream_t saved = c.stream;
  if (stream) c.stream = (hipStream_t)stream;
  int rc = 0;
  const Dims &d = c.d;
  const int npz = c.npz;
  const long long nt = c.ntiles;
  std::vector<int> kt(nq, kord_tr);
  if (form == FV3_FORM_PLAIN64) {
    // working copies: the plain form runs the in-place Fortran semantics
    const size_t bq = sizeof(double) * d.s2 * npz * nq * nt, bdp = sizeof(double) * d.s2 * npz * nt;
    const size_t bcx = sizeof(double) * d.sx * npz * nt, bcy = sizeof(double) * d.sy * npz * nt;
    const size_t bmx = sizeof(double) * d.smx * npz * nt, bmy = sizeof(double) * d.smy * npz * nt;
    char *w = (char *)ctx_stage(c, bdp + bcx + bcy + bmx + bmy + 1024);
    if (!w) {
      rc = 1;
    } else {
      double *wdp = (double *)w, *wcx = (double *)(w + bdp), *wcy = (double *)(w + bdp + bcx);
      double *wmx = (double *)(w + bdp + bcx + bcy), *wmy = (double *)(w + bdp + bcx + bcy + bmx);
      hipMemcpyAsync(q_out, q_in, bq, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(wdp, dp1, bdp, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(wcx, cx, bcx, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(wcy, cy, bcy, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(wmx, mfx, bmx, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(wmy, mfy, bmy, hipMemcpyDeviceToDevice, c.stream);
      hipMemcpyAsync(delp_out, dp1, bdp, hipMemcpyDeviceToDevice, c.stream);
      // the context's sub-cycling rule (fv3_ctx_set_transport), as the fused forms
      rc = tracer_2d_plain<double>(c, c.step_per_level, (double *)q_out, wdp, wmx, wmy, wcx, wcy, nq, hord,
                                   c.step_per_level ? 0 : c.step_q_split, 1.0, 0, 0, 0.0, 1.0, nullptr);
      if (!rc)
        rc = remap_tracers_plain<double>(c, npz, nq, ak, bk, ptop, delp_out, (double *)q_out,
                                         (double *)pt, kt.data(), kord_tm, t_min, 0);
    }
  } else if (form == FV3_FORM_FUSED64 || form == FV3_FORM_FUSED32) {
    rc = fused_step(c, form == FV3_FORM_FUSED32, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy,
                    ak, bk, ptop, hord, kord_tr, kord_tm, t_min, nsplit_cap);
  } else {
    rc = set_error("unknown form %d", form);
  }
  c.stream = saved;
  return rc;
}

int fv3_ctx_set_timing(fv3_ctx *h, int on) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (on && !c.ev[0])
    for (auto &e : c.ev) FV3_HIP(hipEventCreate(&e));
  c.timing = on != 0;
  c.ev_valid = false;
  return 0;
}

int fv3_step_timing(fv3_ctx *h, double *transport_ms, double *remap_ms, int *passes) {
  if (!h || !transport_ms || !remap_ms || !passes) return set_error("fv3_step_timing: NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.ev_valid) return set_error("no timed fused step recorded (fv3_ctx_set_timing)");
  float a = 0.f, b = 0.f;
  FV3_HIP(hipEventSynchronize(c.ev[3]));
  FV3_HIP(hipEventElapsedTime(&a, c.ev[0], c.ev[1]));
  FV3_HIP(hipEventElapsedTime(&b, c.ev[2], c.ev[3]));
  *transport_ms = a;
  *remap_ms = b;
  *passes = c.last_passes;
  return 0;
}

int fv3_ctx_set_transport(fv3_ctx *h, int per_level, int q_split) {
  if (!h) return set_error("NULL context");
  if (q_split < 0) return set_error("q_split must be >= 0 (got %d)", q_split);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  c.step_per_level = per_level != 0;
  c.step_q_split = q_split;
  return 0;
}

int fv3_ctx_set_concurrency(fv3_ctx *h, int on) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  c.concurrent_kinds = on != 0;
  return 0;
}

int fv3_ctx_set_pipeline(fv3_ctx *h, int tracers_per_chunk) {
  if (!h) return set_error("NULL context");
  if (tracers_per_chunk < -1)
    return set_error("tracers per chunk must be -1 (one round), 0 (off) or positive (got %d)", tracers_per_chunk);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  c.pipe_chunk = tracers_per_chunk;
  return 0;
}

int fv3_ctx_set_launch(fv3_ctx *h, int cfg, int rows, int remap) {
  if (!h) return set_error("NULL context");
  if (cfg < 0 || cfg >= FUSED32_NCFG) return set_error("launch configuration %d out of 0..%d", cfg, FUSED32_NCFG - 1);
  if (rows < 0 || (rows > 0 && rows < 6)) return set_error("rows per block must be 0 (default) or >= 6 (got %d)", rows);
  if (remap < 0 || remap > 2) return set_error("remap kernel %d out of 0..2", remap);
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  c.fused_cfg = cfg;
  c.fused_h = rows;
  c.remap_mode = remap;
  return 0;
}

// ---- tangent-linear / adjoint of the FP64 step --------------------------------
static int tlad_args(fv3_ctx *h, int nq, const void *a, const void *b, const void *cc, const double *dp1,
                     const double *cx, const double *cy, const double *mfx, const double *mfy,
                     const double *ak, const double *bk, int kord_tr) {
  if (!h) return set_error("NULL context");
  if (!h->c.grid_set) return set_error("grid metrics not uploaded (fv3_ctx_set_grid)");
  if (!a || !b || !cc || !dp1 || !cx || !cy || !mfx || !mfy) return set_error("NULL argument");
  if (kord_tr && (!ak || !bk)) return set_error("NULL ak/bk with the remap enabled");
  if (nq < 1) return set_error("nq=%d < 1", nq);
  return 0;
}

int fv3_step_tl_dev(fv3_ctx *h, int per_level, int q_split, int nq, const double *q0,
                    const double *dq0, double *q1, double *dq1, const double *dp1,
                    double *delp_out, const double *cx, const double *cy, const double *mfx,
                    const double *mfy, const double *ak, const double *bk, double ptop, int hord,
                    int kord_tr, void *stream) {
  if (tlad_args(h, nq, q0, dq0, q1, dp1, cx, cy, mfx, mfy, ak, bk, kord_tr)) return 1;
  if (!dq1) return set_error("NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t saved = c.stream;
  if (stream) c.stream = (hipStream_t)stream;
  const int rc = step_tl(c, per_level != 0, q_split, nq, q0, dq0, q1, dq1, dp1, delp_out, cx, cy, mfx, mfy,
                         ak, bk, ptop, hord, kord_tr);
  c.stream = saved;
  return rc;
}

int fv3_step_ad_dev(fv3_ctx *h, int per_level, int q_split, int nq, const double *q0,
                    const double *qbar1, double *qbar0, const double *dp1, const double *cx,
                    const double *cy, const double *mfx, const double *mfy, const double *ak,
                    const double *bk, double ptop, int hord, int kord_tr, void *stream) {
  if (tlad_args(h, nq, q0, qbar1, qbar0, dp1, cx, cy, mfx, mfy, ak, bk, kord_tr)) return 1;
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  hipStream_t saved = c.stream;
  if (stream) c.stream = (hipStream_t)stream;
  const int rc = step_ad(c, per_level != 0, q_split, nq, q0, qbar1, qbar0, dp1, cx, cy, mfx, mfy, ak, bk,
                         ptop, hord, kord_tr);
  c.stream = saved;
  return rc;
}

// ---- graph-safe steps --------------------------------------------------------
int fv3_ctx_set_graph_safe(fv3_ctx *h, int on) {
  if (!h) return set_error("NULL context");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (on && !c.step_flag) {
    FV3_HIP(hipMalloc(&c.step_flag, sizeof(int)));
    FV3_HIP(hipMemset(c.step_flag, 0, sizeof(int)));
  }
  c.graph_safe = on != 0;
  return 0;
}

int fv3_step_status(fv3_ctx *h, int *nsplit_overflow) {
  if (!h || !nsplit_overflow) return set_error("fv3_step_status: NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  if (!c.step_flag) return set_error("fv3_step_status: graph-safe mode was never enabled");
  // the step (or a replay of it) may have run on any stream of the device:
  // wait for all of them before reading its flag
  FV3_HIP(hipSetDevice(c.dev));
  FV3_HIP(hipDeviceSynchronize());
  int v = 0;
  FV3_HIP(hipMemcpy(&v, c.step_flag, sizeof(int), hipMemcpyDeviceToHost));
  *nsplit_overflow = v;
  return 0;
}

}  // extern "C"
