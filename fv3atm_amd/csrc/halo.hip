// halo.hip -- the one-device equivalent of FMS mpp_update_domains on the
// six faces of a cubed sphere: scalar fields at cell centres, and vector
// pairs on the A (both components at cell centres), C (u on x-faces, v on
// y-faces) and D (u on y-faces, v on x-faces) staggerings, with the
// components turned into the receiving face's index directions where the
// neighbour face is rotated (the DGRID_NE / CGRID_NE / AGRID flags of the
// FMS call; moving_nest/fv_moving_nest.F90:355-387 shows the fields it
// serves).
//
// Geometry.  A face's points live on a lattice in doubled index units: cell
// centres at odd coordinates, corners at even ones, the face spanning
// [-N, N] in each; a halo point overshoots one face edge.  Folding it across
// that edge (a quarter turn about the edge) lands on the neighbour face,
// whose own coordinates give the source point and whose axes give the
// component rotation: the receiving face's +i direction maps onto one of
// the neighbour's +-i, +-j directions.  The face frames follow the FMS cubic
// mosaic (the orientation the scalar halo table of capi.hip encodes).
// Points that overshoot two edges (the corner blocks) exist on no face and
// are left alone, as FMS does.
#include <algorithm>
#include <array>
#include <cstdlib>
#include <vector>

#include "../../include/fv3atm_amd.h"
#include "fv3_internal.h"

namespace fv3 {

namespace {

// unit normal, +i and +j axes of each face (FMS cubic mosaic orientation)
constexpr int FN[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {-1, 0, 0}, {0, -1, 0}, {0, 0, -1}};
constexpr int FX[6][3] = {{0, 1, 0}, {-1, 0, 0}, {-1, 0, 0}, {0, 0, -1}, {0, 0, -1}, {0, 1, 0}};
constexpr int FY[6][3] = {{0, 0, 1}, {0, 0, 1}, {0, -1, 0}, {0, -1, 0}, {1, 0, 0}, {1, 0, 0}};

using V3 = std::array<int, 3>;
int dot(const V3 &a, const int *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

// Folds the doubled-unit point (I, J) of face f (0-based) onto the face it
// lies on.  Returns false for points on face f itself or in a corner block.
// On success: the source face, its doubled coordinates (Is, Js), and the
// images of the receiving face's +I and +J directions in the source's
// (I, J) axes.
struct Fold {
  int face, I, J;
  int xi, xj;  // image of +I: xi * (+I_s) + xj * (+J_s), one of them +-1
  int yi, yj;  // image of +J
};

bool fold(int n, int f, int I, int J, Fold &out) {
  V3 L, ex, ey;
  for (int c = 0; c < 3; c++) {
    L[c] = n * FN[f][c] + I * FX[f][c] + J * FY[f][c];
    ex[c] = FX[f][c];
    ey[c] = FY[f][c];
  }
  int over = 0, oc = -1;
  for (int c = 0; c < 3; c++)
    if (std::abs(L[c]) > n) {
      over++;
      oc = c;
    }
  if (over != 1) return false;
  const int nc = FN[f][0] != 0 ? 0 : (FN[f][1] != 0 ? 1 : 2);
  const int s = L[oc] > 0 ? 1 : -1, t = L[nc] > 0 ? 1 : -1;
  const int d = std::abs(L[oc]) - n;
  // the quarter turn about the edge: a step outward along oc becomes a step
  // along -t e_nc (toward the neighbour face's interior)
  auto turn = [&](V3 v) {
    V3 r = v;
    r[oc] = 0;
    r[nc] = v[nc] - t * s * v[oc];  // v[nc] is 0 for in-plane directions of face f
    return r;
  };
  L[oc] = s * n;
  L[nc] = t * (n - d);
  int face = -1;
  for (int g = 0; g < 6; g++)
    if (FN[g][oc] == s && FN[g][(oc + 1) % 3] == 0 && FN[g][(oc + 2) % 3] == 0) face = g;
  if (face < 0) return false;
  const V3 dx = turn(ex), dy = turn(ey);
  out.face = face;
  out.I = dot(L, FX[face]);
  out.J = dot(L, FY[face]);
  out.xi = dot(dx, FX[face]);
  out.xj = dot(dx, FY[face]);
  out.yi = dot(dy, FX[face]);
  out.yj = dot(dy, FY[face]);
  return true;
}

// point kinds in doubled units: centre (odd, odd), y-face (odd, even),
// x-face (even, odd); their arrays' extents and offsets
struct Shape {
  int i0, j0, ni, nj;  // lower index bounds and extents
  int ox, oy;          // I = 2 i - ox - N, J = 2 j - oy - N
};

Shape shape_of(int kind, const Dims &d) {
  switch (kind) {
    case 1:  // y-face (south edge of cell (i,j)): (isd:ied, jsd:jed+1)
      return Shape{d.isd, d.jsd, d.nid, d.njd + 1, 1, 2};
    case 2:  // x-face (west edge of cell (i,j)): (isd:ied+1, jsd:jed)
      return Shape{d.isd, d.jsd, d.nid + 1, d.njd, 2, 1};
    default:  // centre: (isd:ied, jsd:jed)
      return Shape{d.isd, d.jsd, d.nid, d.njd, 1, 1};
  }
}

// (the parity of the offset from the face edge, I + N, tells the kind)
int kind_of(int n, int I, int J) { return ((I + n) & 1) ? (((J + n) & 1) ? 0 : 1) : 2; }

}  // namespace

// One halo entry: destination (array, offset within a tile's slab), source
// (tile, array, offset), sign.  Arrays: 0 = the first field (scalar or u),
// 1 = the second (v).
struct HaloEntry {
  int dtile, darr, soff_tile, sarr;
  long long doff, soff;
  float sign;
};

// Builds the table of a six-face context: grid 0 scalar, FV3_AGRID,
// FV3_CGRID, FV3_DGRID.
static std::vector<HaloEntry> build_table(const Dims &d, int grid) {
  const int n = d.ie - d.is + 1;
  std::vector<HaloEntry> out;
  // the point kind of each array: scalar / A: centres; C: u x-faces, v
  // y-faces; D: u y-faces, v x-faces
  int kinds[2] = {0, 0};
  const int narr = grid == 0 ? 1 : 2;
  if (grid == FV3_CGRID) kinds[0] = 2, kinds[1] = 1;
  if (grid == FV3_DGRID) kinds[0] = 1, kinds[1] = 2;
  for (int t = 0; t < 6; t++)
    for (int a = 0; a < narr; a++) {
      const Shape sh = shape_of(kinds[a], d);
      for (int jj = 0; jj < sh.nj; jj++)
        for (int ii = 0; ii < sh.ni; ii++) {
          const int i = sh.i0 + ii, j = sh.j0 + jj;
          const int I = 2 * i - sh.ox - n, J = 2 * j - sh.oy - n;
          Fold fo;
          if (!fold(n, t, I, J, fo)) continue;
          // the component this array holds points along +i (u: a = 0) or +j
          // (v: a = 1) of the receiving face; its image is along the
          // source's +-i or +-j, which selects the source array
          int ci, cj;
          if (narr == 1) {
            ci = 1;
            cj = 0;
          } else if (a == 0) {
            ci = fo.xi;
            cj = fo.xj;
          } else {
            ci = fo.yi;
            cj = fo.yj;
          }
          const int skind = kind_of(n, fo.I, fo.J);
          int sarr = 0;
          if (narr == 2) sarr = ci != 0 ? 0 : 1;
          const int want = narr == 1 ? 0 : kinds[sarr];
          if (skind != want) continue;  // cannot happen on a consistent mosaic
          const Shape ss = shape_of(skind, d);
          const int si = (fo.I + ss.ox + n) / 2, sj = (fo.J + ss.oy + n) / 2;
          HaloEntry e;
          e.dtile = t;
          e.darr = a;
          e.doff = (long long)jj * sh.ni + ii;
          e.soff_tile = fo.face;
          e.sarr = sarr;
          e.soff = (long long)(sj - ss.j0) * ss.ni + (si - ss.i0);
          e.sign = narr == 1 ? 1.f : (float)(ci != 0 ? ci : cj);
          out.push_back(e);
        }
    }
  return out;
}

template <typename T>
__global__ void k_halo_apply(const HaloEntry *tab, int nent, int nslab, T *a0, T *a1, long long s0,
                             long long s1, long long t0, long long t1) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nent * nslab) return;
  const int e = (int)(idx % nent), k = (int)(idx / nent);
  const HaloEntry h = tab[e];
  T *src = h.sarr ? a1 : a0;
  T *dst = h.darr ? a1 : a0;
  const long long ss = h.sarr ? s1 : s0, ds = h.darr ? s1 : s0;
  const long long st = h.sarr ? t1 : t0, dt = h.darr ? t1 : t0;
  dst[h.dtile * dt + k * ds + h.doff] = T(h.sign) * src[h.soff_tile * st + k * ss + h.soff];
}

// Halo tables are built on first use per grid kind and kept in the context.
static int halo_table(Ctx &c, int grid, const HaloEntry **tab, int *n) {
  if (!c.halo_tab[grid]) {
    if (c.graph_safe)
      return set_error("halo table of grid kind %d not built yet (graph-safe mode allocates nothing)", grid);
    const std::vector<HaloEntry> t = build_table(c.d, grid);
    void *p = nullptr;
    FV3_HIP(hipMalloc(&p, sizeof(HaloEntry) * std::max<size_t>(1, t.size())));
    if (!t.empty()) FV3_HIP(hipMemcpy(p, t.data(), sizeof(HaloEntry) * t.size(), hipMemcpyHostToDevice));
    c.halo_tab[grid] = p;
    c.halo_tab_n[grid] = (int)t.size();
  }
  *tab = (const HaloEntry *)c.halo_tab[grid];
  *n = c.halo_tab_n[grid];
  return 0;
}

static int halo_run(Ctx &c, int grid, void *a0, void *a1, int nslab, int elem, hipStream_t s) {
  if (c.ntiles != 6) return set_error("halo updates need a six-tile context");
  if (nslab < 1) return set_error("nslab must be >= 1 (got %d)", nslab);
  if (elem != 4 && elem != 8) return set_error("element size must be 4 or 8 bytes (got %d)", elem);
  const HaloEntry *tab = nullptr;
  int nent = 0;
  if (halo_table(c, grid, &tab, &nent)) return 1;
  int k0 = 0, k1 = 0;
  if (grid == FV3_CGRID) k0 = 2, k1 = 1;
  if (grid == FV3_DGRID) k0 = 1, k1 = 2;
  const Shape h0 = shape_of(k0, c.d), h1 = shape_of(k1, c.d);
  const long long s0 = (long long)h0.ni * h0.nj, s1 = (long long)h1.ni * h1.nj;
  const long long total = (long long)nent * nslab;
  if (total == 0) return 0;
  const dim3 b(256), g((unsigned)((total + 255) / 256));
  if (elem == 8)
    hipLaunchKernelGGL(k_halo_apply<double>, g, b, 0, s, tab, nent, nslab, (double *)a0, (double *)a1, s0,
                       s1, s0 * nslab, s1 * nslab);
  else
    hipLaunchKernelGGL(k_halo_apply<float>, g, b, 0, s, tab, nent, nslab, (float *)a0, (float *)a1, s0, s1,
                       s0 * nslab, s1 * nslab);
  FV3_HIP(hipGetLastError());
  return 0;
}

}  // namespace fv3

using namespace fv3;

extern "C" {

int fv3_halo_update(fv3_ctx *h, void *field, int nslab, int elem_bytes, void *stream) {
  if (!h || !field) return set_error("fv3_halo_update: NULL argument");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  return halo_run(c, 0, field, field, nslab, elem_bytes, stream ? (hipStream_t)stream : c.stream);
}

int fv3_halo_update_vector(fv3_ctx *h, void *u, void *v, int nslab, int grid, int elem_bytes,
                           void *stream) {
  if (!h || !u || !v) return set_error("fv3_halo_update_vector: NULL argument");
  if (grid != FV3_AGRID && grid != FV3_CGRID && grid != FV3_DGRID)
    return set_error("grid must be FV3_AGRID, FV3_CGRID or FV3_DGRID (got %d)", grid);
  if (u == v) return set_error("fv3_halo_update_vector: u and v must be distinct arrays");
  Ctx &c = h->c;
  std::lock_guard<std::mutex> lk(c.mu);
  return halo_run(c, grid, u, v, nslab, elem_bytes, stream ? (hipStream_t)stream : c.stream);
}

}  // extern "C"
