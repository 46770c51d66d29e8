// fused_kernel.h -- the fused transport kernel (one launch per sub-cycle
// pass over every level whose nsplit(k) reaches the pass), instantiated for
// FP64 (fused.hip, contraction off) and FP32 (fused32.hip).
#pragma once
#include "fused_tp.h"

namespace fv3 {

// In/out buffers of level k at pass `it`: the last sub-cycle of every level
// lands in q_out; earlier ones alternate through q_tmp.
template <typename T>
__device__ __forceinline__ void pass_buffers(const FusedArgs<T> &a, int nspl, const T *&in, T *&out) {
  in = a.it == 1 ? a.q_in : (((nspl - (a.it - 1)) & 1) == 0 ? a.q_out : a.q_tmp);
  out = ((nspl - a.it) & 1) == 0 ? a.q_out : a.q_tmp;
}

// per (strip, block, level) constants of a wave
template <typename T>
struct StripCtx {
  int i, ic, j0, h, nix, niy, lo_q, lo_fx, lo_mx;  // lane column; block rows; lane offsets in the row layouts
  bool col_out, lane_in, edge_strip, edge_rows, corner_wg, scale;
  bool gen_src;  // q sources by the general (halo, six-tile, corner) rules
  T frac;
  const T *cxk, *cyk, *mfxk, *mfyk;
  const double *dp1k;
  long long tg, tdx, tdy, tgf;  // this tile's offsets in the metric arrays (s2, dx, dy, gfac)
  EdgeCoef ec;                  // FP32 edge strips: the lane's cube-edge coefficients
};

// per tracer streaming state
template <typename T>
struct TrState {
  YPipe<T> py, pj;
  T fx2_r[6];
  T fy2_prev, Fy_prev;
  const T *qin, *qhalo;
  const T *qsrc0;  // pass input of tile 1 at this tracer and level (six-tile halos)
  T *qout;
  __amdgpu_buffer_rsrc_t qout_r;  // FP32: qout as a buffer (scalar row offset, lane offset)
  __amdgpu_buffer_rsrc_t qin_r;   // FP32: qin likewise
};

// level data carried between rows
template <typename T>
struct LvState {
  T area_r[6], cx_r[6];
  T yfx_prev, mfy_prev;
};

// Everything row_step reads from HBM for one input row, fetched one row ahead
// so that the loads of row t+1 are in flight while row t computes.  Upwind
// metric candidates for both signs of the Courant number are fetched, so no
// address depends on a loaded value.
template <typename T, int NT>
struct RowIn {
  T area, cyu, cxu, mfy, mfx, rarea;
  T ya, yb, yc, yd, ye;  // FP32: gyp gym; FP64: dya(j-1) sg3(j-1) dya(j) sg1(j) dx
  T xa, xb, xc, xd, xe;  // FP32: gxp gxm; FP64: dxa(i-1) sg2(i-1) dxa(i) sg0(i) dy
  double dp1;
  T area_m, cxm_u;  // FP32: area and cx of row r-3, from the LDS ring
  T qy[NT], qx[NT];
  T qw[NT][4];      // FP32: the y-sweep sources of rows r-1 .. r-4, from the LDS ring
  T fx2m[NT];       // FP32: the inner x-flux of row r-3 (kept in the ring by this wave)
  float *fx2w[NT];  // FP32: where this row's inner x-flux goes
};

// q sources per input row of a kind: the strips with corner cells read the
// y- and x-sweep sources of a cell separately
constexpr int ek_qn(int ek) { return ek >= 2 && ek <= 4 ? 2 : 1; }
// x-sweep cube-edge side of a kind: 1 left only, 2 right only, 0 either
constexpr int ek_side(int ek) { return ek == 3 || ek == 5 ? 1 : (ek == 4 || ek == 6 ? 2 : 0); }

template <typename T>
struct RowIdx {
  int r, rc, m, mc, fc;
  long long orow, corow, co;  // row offsets are wave-uniform; co adds the lane
  bool row_in, out_row;
};

template <typename T>
__device__ __forceinline__ RowIdx<T> row_idx(const FusedArgs<T> &a, const StripCtx<T> &sc, int t) {
  RowIdx<T> x;
  x.r = sc.j0 - NG + t;
  x.rc = x.r > a.jed ? a.jed : x.r;
  x.m = x.r - 3;
  x.mc = x.m < a.js ? a.js : (x.m > a.je ? a.je : x.m);
  const int f = x.r - 2;
  x.fc = f < a.js ? a.js : (f > a.je + 1 ? a.je + 1 : f);
  x.orow = (long long)(x.rc - a.jsd) * a.nid;
  x.row_in = x.rc >= a.js && x.rc <= a.je;
  x.out_row = t >= 2 * NG && x.m >= sc.j0 && x.m < sc.j0 + sc.h && x.m <= a.je;
  x.corow = (long long)(x.mc - a.jsd) * a.nid;
  x.co = x.corow + sc.lo_q;
  return x;
}

// Lane source of the pass input q at input row x for the y-sweep (py) and
// the x-sweep (px): the cell itself inside the compute domain; on a frame
// block, halo cells of one tile come from the caller's halo, halo cells of
// six tiles from the neighbour face, and corner cells from their
// copy_corners source (dir 2 for the y-sweep, dir 1 for the x-sweep).
template <typename T, int EK>
__device__ __forceinline__ void q_src(const FusedArgs<T> &a, const StripCtx<T> &sc, const TrState<T> &s,
                                      const RowIdx<T> &x, const T *&py, const T *&px) {
  const int nid = a.nid;
  // EK 1 blocks have every lane inside the compute domain, EK 0 every cell
  const bool inside = EK == 0 || ((EK == 1 || sc.lane_in) && x.row_in);  // EK >= 2: strips
  if constexpr (EK >= 5) {  // edge strips, lean rows (inside the domain): halo lanes, no corner cell
    py = (sc.lane_in ? s.qin : s.qhalo) + x.orow + sc.lo_q;
    if (a.ntiles > 1 && !sc.lane_in) {
      int ts, si, sj;
      cube_halo_src(blockIdx.z + 1, sc.ic, x.rc, a.ie, ts, si, sj);
      py = s.qsrc0 + (long long)(ts - 1) * a.tq + (long long)(sj - a.jsd) * nid + (si - a.isd);
    }
    px = py;
    return;
  }
  py = (inside ? s.qin : s.qhalo) + x.orow + sc.lo_q;
  px = py;
  if constexpr (EK == 1) {
    if (a.ntiles > 1 && !inside) {  // edge halo rows of a cube face (no corner cell)
      int ts, si, sj;
      cube_halo_src(blockIdx.z + 1, sc.ic, x.rc, a.ie, ts, si, sj);
      py = s.qsrc0 + (long long)(ts - 1) * a.tq + (long long)(sj - a.jsd) * nid + (si - a.isd);
      px = py;
    }
  }
  if constexpr (EK >= 2 && EK <= 4) {
    if (a.ntiles > 1 && !inside) {
      const int n = a.ie;
      const int tile = blockIdx.z + 1;
      auto from_cube = [&](int ci, int cj) -> const T * {
        int ts, si, sj;
        cube_halo_src(tile, ci, cj, n, ts, si, sj);
        return s.qsrc0 + (long long)(ts - 1) * a.tq + (long long)(sj - a.jsd) * nid + (si - a.isd);
      };
      int si, sj;
      if (corner_src(sc.ic, x.rc, a.npx, a.npy, 2, a.corners, si, sj)) {
        py = from_cube(si, sj);
        corner_src(sc.ic, x.rc, a.npx, a.npy, 1, a.corners, si, sj);
        px = from_cube(si, sj);
      } else {
        py = from_cube(sc.ic, x.rc);
        px = py;
      }
    } else if (sc.corner_wg && !x.row_in && !sc.lane_in) {
      int si, sj;
      if (corner_src(sc.ic, x.rc, a.npx, a.npy, 2, a.corners, si, sj)) {
        py = s.qhalo + (long long)(sj - a.jsd) * nid + (si - a.isd);
        corner_src(sc.ic, x.rc, a.npx, a.npy, 1, a.corners, si, sj);
        px = s.qhalo + (long long)(sj - a.jsd) * nid + (si - a.isd);
      }
    }
  }
}

template <typename T, int NT>
__device__ __forceinline__ void row_fetch(const FusedArgs<T> &a, const StripCtx<T> &sc,
                                          const TrState<T> (&st)[NT], const bool (&act)[NT], int t,
                                          RowIn<T, NT> &in) {
  constexpr bool fp64 = sizeof(T) == 8;
  const RowIdx<T> x = row_idx(a, sc, t);
  const int nid = a.nid;
  // wave-uniform row bases, indexed by the lane's 32-bit column offset
  in.area = (a.area + sc.tg + x.orow)[sc.lo_q];
  in.cyu = (sc.cyk + (long long)(x.fc - a.js) * nid)[sc.lo_q];
  const long long frow = (long long)(x.fc - a.jsd) * nid;
  const long long gxrow = (long long)(x.rc - a.jsd) * (nid + 1) + (a.is - a.isd);
  const long long urow = x.orow + (a.is - a.isd);
  if constexpr (fp64) {
    in.ya = (a.dya + sc.tg + frow - nid)[sc.lo_q];
    in.yb = (a.sin_sg + 9 * sc.tg + 3 * a.ms2 + frow - nid)[sc.lo_q];
    in.yc = (a.dya + sc.tg + frow)[sc.lo_q];
    in.yd = (a.sin_sg + 9 * sc.tg + 1 * a.ms2 + frow)[sc.lo_q];
    in.ye = (a.dx + sc.tdx + frow)[sc.lo_q];
    in.xa = (a.dxa + sc.tg + urow - 1)[sc.lo_fx];
    in.xb = (a.sin_sg + 9 * sc.tg + 2 * a.ms2 + urow - 1)[sc.lo_fx];
    in.xc = (a.dxa + sc.tg + urow)[sc.lo_fx];
    in.xd = (a.sin_sg + 9 * sc.tg + 0 * a.ms2 + urow)[sc.lo_fx];
    in.xe = (a.dy + sc.tdy + gxrow)[sc.lo_fx];
  } else {
    in.ya = (a.gyp + sc.tgf + frow)[sc.lo_q];
    in.yb = (a.gym + sc.tgf + frow)[sc.lo_q];
    in.xa = (a.gxp + sc.tgf + gxrow)[sc.lo_fx];
    in.xb = (a.gxm + sc.tgf + gxrow)[sc.lo_fx];
  }
  in.cxu = (sc.cxk + (long long)(x.rc - a.jsd) * sc.nix)[sc.lo_fx];
  in.mfy = (sc.mfyk + (long long)(x.fc - a.js) * sc.niy)[sc.lo_mx];
  in.mfx = (sc.mfxk + (long long)(x.mc - a.js) * sc.nix)[sc.lo_fx];
  in.rarea = T(0);
  in.dp1 = 0.0;
  if (x.out_row) {
    in.rarea = (a.rarea + sc.tg + x.corow)[sc.lo_q];
    in.dp1 = (sc.dp1k + x.corow)[sc.lo_q];
  }
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (!act[n]) continue;
    const T *py, *px;
    if (sc.gen_src)
      q_src<T, 2>(a, sc, st[n], x, py, px);
    else
      q_src<T, 0>(a, sc, st[n], x, py, px);
    const T qy = *py;
    const T qx = sc.gen_src ? *px : qy;
    in.qy[n] = qy;
    in.qx[n] = qx;
  }
}

// One input row r = j0 - 3 + t of the streaming pass; U = t mod 6.  HO: the
// advection order when fixed at compile time (0: taken from the arguments).
// TW = t for the block's first six input rows, 6 after them.  Those rows
// (r < j0 + 3) feed only the inner x-sweep, q_j and the two y-pipes: the
// q_i, the outer x-sweep and the update of row r-3 < j0 are never used, and
// the y-pipes' stages start as their inputs fill (first needed face: j0, at
// t = 5).
template <typename T, int NT, int U, int HO, int SIDE = 0, int TW = 6>
__device__ __forceinline__ void row_step(const FusedArgs<T> &a, const StripCtx<T> &sc,
                                         LvState<T> &lv, TrState<T> (&st)[NT], const bool (&act)[NT],
                                         int t, const RowIn<T, NT> &in) {
  constexpr bool fp64 = sizeof(T) == 8;
  constexpr int M3 = (U + 3) % 6;  // slot of row r-3
  const int nid = a.nid;
  const RowIdx<T> x = row_idx(a, sc, t);
  const int r = x.r;
  const int hord = HO ? HO : a.hord;
  const int ord_in = hord == 10 ? 8 : hord;

  // ---- level data of this row (shared by the NT tracers)
  const T area_r = in.area;
  if constexpr (fp64) lv.area_r[U] = area_r;
  const T cyu = in.cyu;
  // x * (1/nsplit) with 1/1 = 1 exactly: unconditional scaling rounds the same
  const T cys = TW >= 5 ? cyu * sc.frac : T(0);
  T yfx = T(0);
  if constexpr (TW >= 5) {
    const bool pos = cyu > T(0);
    if constexpr (fp64)
      yfx = pos ? cyu * in.ya * in.ye * in.yb : cyu * in.yc * in.ye * in.yd;
    else
      yfx = cyu * (pos ? in.ya : in.yb);
    yfx = yfx * sc.frac;
  }
  const T cxr_u = in.cxu;
  if constexpr (fp64) lv.cx_r[U] = cxr_u;
  const T cxr = cxr_u * sc.frac;
  T xfx;
  {
    const bool pos = cxr_u > T(0);
    if constexpr (fp64)
      xfx = pos ? cxr_u * in.xa * in.xe * in.xb : cxr_u * in.xc * in.xe * in.xd;
    else
      xfx = cxr_u * (pos ? in.xa : in.xb);
    xfx = xfx * sc.frac;
  }
  const T xfxp = from_right(xfx);
  const T rax = area_r + xfx - xfxp;
  const T area_m = fp64 ? lv.area_r[M3] : in.area_m;
  T rrax = T(0), rray = T(0), ray = T(1);
  if constexpr (TW >= 6) ray = area_m + lv.yfx_prev - yfx;
  if constexpr (!fp64) {
    rrax = __builtin_amdgcn_rcpf(rax);
    if constexpr (TW >= 6) rray = __builtin_amdgcn_rcpf(ray);
  }
  const T cxm_u = fp64 ? lv.cx_r[M3] : in.cxm_u;  // cx of row m (row r-3, loaded 3 steps ago)
  const T cxm = TW >= 6 ? cxm_u * sc.frac : T(0);
  const T mfys = TW >= 5 ? in.mfy * sc.frac : T(0);
  const T mfxs = TW >= 6 ? in.mfx * sc.frac : T(0);
  const T mfxp = TW >= 6 ? from_right(mfxs) : T(0);
  // edge strips: dxa of the rows (FP64), or the cells' edge face weights (FP32)
  T dxa_m = T(1), dxa_r = T(1);
  if (sc.edge_strip) {
    const T *g = fp64 ? a.dxa + sc.tg : a.gex + sc.tgf;
    if constexpr (TW >= 6) dxa_m = g[(long long)(x.mc - a.jsd) * nid + sc.lo_q];
    dxa_r = g[x.orow + sc.lo_q];
  }
  // cube-edge role of cell r-2 along the columns (uniform across the wave)
  const int role_y = TW >= 4 && sc.edge_rows ? edge_role(r - 2, a.js, a.je, a.npy, a.edge) : 0;
  T dd0 = T(0), dd1 = T(0), dd2 = T(0), dd3 = T(0), dd4 = T(0);
  if (role_y != 0) {
    auto dya_at = [&](int jj) {
      jj = jj < a.jsd ? a.jsd : (jj > a.jed ? a.jed : jj);
      return (a.dya + sc.tg)[(long long)(jj - a.jsd) * nid + sc.lo_q];
    };
    dd0 = dya_at(r - 4);
    dd1 = dya_at(r - 3);
    dd2 = dya_at(r - 2);
    dd3 = dya_at(r - 1);
    dd4 = dya_at(r);
  }
  // output row: thickness and area factors shared by the tracers
  // (FP32 reads rarea and dp1 of every row from the ring and computes them
  // without a branch; only output rows store)
  T rarea = T(0), d1 = T(0), rdp2 = T(0), dp2 = T(0);
  if (TW >= 6 && (!fp64 || x.out_row)) {
    rarea = in.rarea;
    const T ddiv = ((mfxs - mfxp) + (lv.mfy_prev - mfys)) * rarea;
    T xd = (T)in.dp1;
    for (int s = 1; s < a.it; s++) xd = xd + ddiv;
    d1 = xd;
    dp2 = xd + ddiv;
    if constexpr (!fp64) rdp2 = __builtin_amdgcn_rcpf(dp2);
  }

  // ---- the tracers
#pragma unroll
  for (int n = 0; n < NT; n++) {
    // FP32: an inactive tracer of an active wave computes on tracer 0's rows
    // (what its ring slots hold) and stores nothing, so no branch per tracer
    if (fp64 && !act[n]) continue;
    TrState<T> &s = st[n];
    const T qy = in.qy[n], qx = in.qx[n];
    // inner y-sweep at face r-2 and q_i of row r-3
    const T fy2 = ypipe_step<T, U, TW, !fp64>(s.py, qy, in.qw[n], cys, ord_in, role_y, dd0, dd1, dd2, dd3, dd4);
    const T qm = fp64 ? s.py.w[M3] : in.qw[n][2];
    // outer x-sweep of row r-3, inner x-sweep of row r, q_j of row r
    T fx = T(0);
    if constexpr (TW >= 6) {
      T q_i;
      if constexpr (fp64)
        q_i = (qm * area_m + lv.yfx_prev * s.fy2_prev - yfx * fy2) / ray;
      else
        q_i = (qm * area_m + lv.yfx_prev * s.fy2_prev - yfx * fy2) * rray;
      fx = xsweep<SIDE>(q_i, cxm, sc.i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, hord, dxa_m, sc.ec);
    }
    const T fx2 = xsweep<SIDE>(qx, cxr, sc.i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, ord_in, dxa_r, sc.ec);
    if constexpr (fp64)
      s.fx2_r[U] = fx2;
    else
      in.fx2w[n][0] = fx2;
    const T fx2p = from_right(fx2);
    T q_j;
    if constexpr (fp64)
      q_j = (qx * area_r + xfx * fx2 - xfxp * fx2p) / rax;
    else
      q_j = (qx * area_r + xfx * fx2 - xfxp * fx2p) * rrax;
    // outer y-sweep at face r-2, fluxes, update of row r-3
    const T fy = ypipe_step<T, U, TW>(s.pj, q_j, nullptr, cys, hord, role_y, dd0, dd1, dd2, dd3, dd4);
    if constexpr (TW < 5) continue;
    const T Fy = T(0.5) * (fy + fy2) * mfys;
    s.fy2_prev = fy2;
    if constexpr (TW < 6) {
      s.Fy_prev = Fy;
      continue;
    }
    const T Fx = T(0.5) * (fx + (fp64 ? s.fx2_r[M3] : in.fx2m[n])) * mfxs;
    const T Fxp = from_right(Fx);
    if (act[n] && x.out_row && sc.col_out) {
      if constexpr (fp64) {
        (s.qout + x.corow)[sc.lo_q] = (qm * d1 + (Fx - Fxp + s.Fy_prev - Fy) * rarea) / dp2;
      } else {  // the row as the scalar offset, the lane's column as the vector offset
        const float v = (qm * d1 + (Fx - Fxp + s.Fy_prev - Fy) * rarea) * rdp2;
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), s.qout_r, sc.lo_q * 4,
                                              (int)x.corow * 4, 0);
      }
    }
    s.Fy_prev = Fy;
  }
  lv.yfx_prev = yfx;
  lv.mfy_prev = mfys;
}

// ---- LDS row ring of the FP32 kernels.  Rows are loaded straight into LDS
// with global_load_lds_dword, or buffer_load_dword ... lds where the row
// address is a scalar offset (the data in flight holds no VGPRs), so the
// loads of a row can be issued LR_AHEAD rows ahead.  The waves of a
// workgroup split the LR_NL level arrays of a row between them (they stream
// the same strip and level) and each loads its own tracers' rows.  The ring
// also keeps the four rows behind the one computed: the stencil's values of
// rows r-1 .. r-4 (y-sweep sources, area and cx of row r-3) are read back
// from LDS instead of being carried in registers, and each wave keeps its
// inner x-fluxes there for the three rows until they are averaged: fewer
// VGPRs, more waves per CU.
constexpr int LR_DEPTH = 8;  // slots: rows t-4 .. t+3 at row t
constexpr int LR_AHEAD = 3;  // rows issued ahead of the row computed
constexpr int LR_NL = 12;  // area cyu cxu mfy mfx rarea gyp gym gxp gxm dp1.lo dp1.hi

template <int W, int NT, int QN>
struct LdsRing {
  // per slot: the level arrays and q sources loaded by the DMA
  float v[LR_DEPTH][LR_NL + W * NT * QN][FT_LANES];
  // each wave's inner x-fluxes of rows t-3 .. t (written and read by the wave itself)
  float fx2[4][W * NT][FT_LANES];
};

__device__ __forceinline__ void lds_dma(const float *src, float *dst_row) {
  __builtin_amdgcn_global_load_lds((const void *)src, (__attribute__((address_space(3))) void *)dst_row, 4, 0, 0);
}

// s_waitcnt vmcnt(n) lgkmcnt(0) (gfx9 encoding: vmcnt in bits 3:0 and 15:14,
// expcnt 6:4 left at its no-wait maximum, lgkmcnt 11:8)
template <int N>
__device__ __forceinline__ void wait_vm_lgkm0() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | ((N >> 4) << 14) | (0x7 << 4));
}

// issue the loads of input row t into ring slot t mod LR_DEPTH
template <int W, int NT, int EK>
__device__ __forceinline__ void ring_issue(const FusedArgs<float> &a, const StripCtx<float> &sc,
                                           const TrState<float> (&st)[NT],
                                           LdsRing<W, NT, ek_qn(EK)> &ring, int wave, int t) {
  constexpr int QN = ek_qn(EK);
  const RowIdx<float> x = row_idx(a, sc, t);
  const int nid = a.nid;
  float(*slot)[FT_LANES] = ring.v[t & (LR_DEPTH - 1)];
  const long long frow = (long long)(x.fc - a.jsd) * nid;
  const long long gxrow = (long long)(x.rc - a.jsd) * (nid + 1) + (a.is - a.isd);
  // uniform row base (scalar) + the lane's column offset in that array's
  // row layout (cell, x-face, y-face column; dp1 as two 32-bit halves)
  for (int j = wave; j < LR_NL; j += W) {
    const float *base;
    int kind;
    switch (j) {
      case 0: base = a.area + sc.tg + x.orow; kind = 0; break;
      case 1: base = sc.cyk + (long long)(x.fc - a.js) * nid; kind = 0; break;
      case 2: base = sc.cxk + (long long)(x.rc - a.jsd) * sc.nix; kind = 1; break;
      case 3: base = sc.mfyk + (long long)(x.fc - a.js) * sc.niy; kind = 2; break;
      case 4: base = sc.mfxk + (long long)(x.mc - a.js) * sc.nix; kind = 1; break;
      case 5: base = a.rarea + sc.tg + x.corow; kind = 0; break;
      case 6: base = a.gyp + sc.tgf + frow; kind = 0; break;
      case 7: base = a.gym + sc.tgf + frow; kind = 0; break;
      case 8: base = a.gxp + sc.tgf + gxrow; kind = 1; break;
      case 9: base = a.gxm + sc.tgf + gxrow; kind = 1; break;
      case 10: base = (const float *)(sc.dp1k + x.corow); kind = 3; break;
      default: base = (const float *)(sc.dp1k + x.corow) + 1; kind = 3; break;
    }
    const int off = kind == 0 ? sc.lo_q : (kind == 1 ? sc.lo_fx : (kind == 2 ? sc.lo_mx : 2 * sc.lo_q));
    lds_dma(base + off, slot[j]);
  }
  // inactive tracers load tracer 0: every wave issues the same count
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const float *py, *px;
    q_src<float, EK>(a, sc, st[n], x, py, px);
    lds_dma(py, slot[LR_NL + (wave * NT + n) * QN]);
    if (QN == 2) lds_dma(px, slot[LR_NL + (wave * NT + n) * QN + 1]);
  }
}

// Lean blocks never clamp a row (every input row and every stencil row lies
// inside the compute domain), so each level array's row address is affine in
// the row: the wave's share of the level arrays is set up once per round as
// (row-0 address, elements per row, lane offset) and each row costs one
// multiply-add per array.  (The edge kinds, short of scalar registers, keep
// the per-row address expressions.)
template <int W>
struct LeanRows {
  static constexpr int M = (LR_NL + W - 1) / W;
  const float *p[M];
  int step[M], loff[M];
  __amdgpu_buffer_rsrc_t r[M];  // p[m] as a buffer: row t at scalar offset 4 t step, lane at 4 loff
};

template <int W>
__device__ __forceinline__ LeanRows<W> lean_rows(const FusedArgs<float> &a, const StripCtx<float> &sc,
                                                 int wave) {
  LeanRows<W> lr;
  const int nid = a.nid;
  const long long r0 = sc.j0 - NG;  // input row of t = 0
#pragma unroll
  for (int m = 0; m < LeanRows<W>::M; m++) {
    const int j = wave + m * W;
    const float *p = a.area;
    int step = nid, loff = sc.lo_q;
    switch (j) {
      case 0: p = a.area + sc.tg + (r0 - a.jsd) * nid; break;
      case 1: p = sc.cyk + (r0 - 2 - a.js) * nid; break;
      case 2: p = sc.cxk + (r0 - a.jsd) * sc.nix; step = sc.nix; loff = sc.lo_fx; break;
      case 3: p = sc.mfyk + (r0 - 2 - a.js) * sc.niy; step = sc.niy; loff = sc.lo_mx; break;
      case 4: p = sc.mfxk + (r0 - 3 - a.js) * sc.nix; step = sc.nix; loff = sc.lo_fx; break;
      case 5: p = a.rarea + sc.tg + (r0 - 3 - a.jsd) * nid; break;
      case 6: p = a.gyp + sc.tgf + (r0 - 2 - a.jsd) * nid; break;
      case 7: p = a.gym + sc.tgf + (r0 - 2 - a.jsd) * nid; break;
      case 8: p = a.gxp + sc.tgf + (r0 - a.jsd) * (nid + 1) + (a.is - a.isd); step = nid + 1; loff = sc.lo_fx; break;
      case 9: p = a.gxm + sc.tgf + (r0 - a.jsd) * (nid + 1) + (a.is - a.isd); step = nid + 1; loff = sc.lo_fx; break;
      case 10: p = (const float *)(sc.dp1k + (r0 - 3 - a.jsd) * nid); step = 2 * nid; loff = 2 * sc.lo_q; break;
      case 11: p = (const float *)(sc.dp1k + (r0 - 3 - a.jsd) * nid) + 1; step = 2 * nid; loff = 2 * sc.lo_q; break;
      default: break;  // past the last array: not issued
    }
    lr.p[m] = p;
    lr.step[m] = step;
    lr.loff[m] = loff;
    lr.r[m] = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(p), (short)0, 0x7fffffff, 0x00020000);
  }
  return lr;
}

template <int W, int NT>
__device__ __forceinline__ void ring_issue_lean(const FusedArgs<float> &a, const StripCtx<float> &sc,
                                                const TrState<float> (&st)[NT], const LeanRows<W> &lr,
                                                LdsRing<W, NT, 1> &ring, int wave, int t) {
  float(*slot)[FT_LANES] = ring.v[t & (LR_DEPTH - 1)];
#pragma unroll
  for (int m = 0; m < LeanRows<W>::M; m++)
    if (wave + m * W < LR_NL)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(lr.r[m], (__attribute__((address_space(3))) void *)slot[wave + m * W], 4,
                                               lr.loff[m] * 4, t * lr.step[m] * 4, 0, 0);
  const int orow = (sc.j0 - NG + t - a.jsd) * a.nid;
#pragma unroll
  for (int n = 0; n < NT; n++)  // inactive tracers load tracer 0: every wave issues the same count
    __builtin_amdgcn_raw_ptr_buffer_load_lds(st[n].qin_r, (__attribute__((address_space(3))) void *)slot[LR_NL + wave * NT + n],
                                             4, sc.lo_q * 4, orow * 4, 0, 0);
}

template <int W, int NT, int EK>
__device__ __forceinline__ void ring_issue_mixed(const FusedArgs<float> &a, const StripCtx<float> &sc,
                                                 const TrState<float> (&st)[NT], const LeanRows<W> &lr,
                                                 LdsRing<W, NT, ek_qn(EK)> &ring, int wave, int t) {
  constexpr int QN = ek_qn(EK);
  float(*slot)[FT_LANES] = ring.v[t & (LR_DEPTH - 1)];
#pragma unroll
  for (int m = 0; m < LeanRows<W>::M; m++)
    if (wave + m * W < LR_NL)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(lr.r[m], (__attribute__((address_space(3))) void *)slot[wave + m * W], 4,
                                               lr.loff[m] * 4, t * lr.step[m] * 4, 0, 0);
  const RowIdx<float> x = row_idx(a, sc, t);
#pragma unroll
  for (int n = 0; n < NT; n++) {
    const float *py, *px;
    q_src<float, EK>(a, sc, st[n], x, py, px);
    lds_dma(py, slot[LR_NL + (wave * NT + n) * QN]);
    if (QN == 2) lds_dma(px, slot[LR_NL + (wave * NT + n) * QN + 1]);
  }
}

template <int W, int NT, int QN>
__device__ __forceinline__ void ring_read(LdsRing<W, NT, QN> &ring, int wave, int lane, int t,
                                          RowIn<float, NT> &in) {
  const float(*slot)[FT_LANES] = ring.v[t & (LR_DEPTH - 1)];
  const float(*slot_m)[FT_LANES] = ring.v[(t - 3) & (LR_DEPTH - 1)];
  in.area_m = slot_m[0][lane];
  in.cxm_u = slot_m[2][lane];
  in.area = slot[0][lane];
  in.cyu = slot[1][lane];
  in.cxu = slot[2][lane];
  in.mfy = slot[3][lane];
  in.mfx = slot[4][lane];
  in.rarea = slot[5][lane];
  in.ya = slot[6][lane];
  in.yb = slot[7][lane];
  in.xa = slot[8][lane];
  in.xb = slot[9][lane];
  const unsigned long long lo = __float_as_uint(slot[10][lane]), hi = __float_as_uint(slot[11][lane]);
  in.dp1 = __longlong_as_double((long long)(lo | (hi << 32)));
#pragma unroll
  for (int n = 0; n < NT; n++) {
#pragma unroll
    for (int d = 1; d <= 4; d++) in.qw[n][d - 1] = ring.v[(t - d) & (LR_DEPTH - 1)][LR_NL + (wave * NT + n) * QN][lane];
    in.fx2m[n] = ring.fx2[(t - 3) & 3][wave * NT + n][lane];
    in.fx2w[n] = &ring.fx2[t & 3][wave * NT + n][lane];
    in.qy[n] = slot[LR_NL + (wave * NT + n) * QN][lane];
    in.qx[n] = QN == 2 ? slot[LR_NL + (wave * NT + n) * QN + 1][lane] : in.qy[n];
  }
}

// wait until this wave's loads of row t are in LDS (loads complete in issue
// order; ahead = rows issued after t; every wave issues at least NMIN loads
// per row), then until every wave's are (barrier)
template <int W, int NT, int QN>
__device__ __forceinline__ void ring_wait(int ahead) {
  constexpr int NMIN = LR_NL / W + NT * QN;
  static_assert((LR_AHEAD - 1) * NMIN < 64, "vmcnt range");
  if (ahead >= 2)
    wait_vm_lgkm0<2 * NMIN>();
  else if (ahead == 1)
    wait_vm_lgkm0<NMIN>();
  else
    wait_vm_lgkm0<0>();
  if (W > 1) __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");  // no LDS access moves across the barrier
}

// EK (edge kind) 0: the lean variant for the interior blocks (no cube-edge,
// corner or halo code is compiled); 1: blocks of the interior strips that
// reach a halo or cube-edge row (no per-lane edge or corner code); 2: the
// strips that reach a halo or cube-edge column when no interior strip
// exists; 3 / 4: the top and bottom blocks of those strips left / right of
// the interior strips (each can only meet its own cube edge); 5 / 6: their
// lean rows (halo lanes and the x-direction cube-edge code, but no corner
// cell, no y-direction edge code and one q source per cell).
// OCC: minimum workgroups per CU asked of the register allocator.
template <typename T, int W, int NT, int HO, int EK, int OCC>
__global__ void __launch_bounds__(FT_LANES *W, OCC) k_fused_tp(const FusedArgs<T> a) {
  const int lane = threadIdx.x & (FT_LANES - 1);
  // wave index as a scalar: everything derived from it stays in SGPRs
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / FT_LANES);
  const int k = blockIdx.y;
  const int nspl = a.nsplit[k];
  // whole block: this level is done, or needs more passes than are launched
  // (reported by the caller's nsplit check; nothing is written for it)
  if (a.it > nspl || nspl > a.cap) return;
  int strip, vb, j0, h;
  {
    const int ns = a.s_hi >= a.s_lo ? a.s_hi - a.s_lo + 1 : 0;
    const int bx = (int)blockIdx.x;
    if constexpr (EK == 0) {  // interior strips, lean rows
      strip = a.s_lo + bx % ns;
      vb = bx / ns;
      j0 = a.lt + vb * a.H;
      h = min(a.H, a.lb + 1 - j0);
    } else if constexpr (EK == 1) {  // interior strips, the rows above and below the lean rows
      strip = a.s_lo + bx % ns;
      vb = bx / ns;
      j0 = vb == 0 ? a.js : a.lb + 1;
      h = vb == 0 ? a.lt - a.js : a.je - a.lb;
    } else if constexpr (EK == 2) {  // every strip (no interior strip exists)
      strip = bx % a.nstrip;
      vb = bx / a.nstrip;
      j0 = a.js + vb * a.H;
      h = min(a.H, a.je + 1 - j0);
    } else {
      // 3 / 5: strips left of the interior strips; 4 / 6: right of them
      const int nside = EK == 3 || EK == 5 ? a.s_lo : a.nstrip - 1 - a.s_hi;
      strip = (EK == 3 || EK == 5 ? 0 : a.s_hi + 1) + bx % nside;
      vb = bx / nside;
      if constexpr (EK <= 4) {  // the top and bottom blocks
        j0 = vb == 0 ? a.js : a.lb + 1;
        h = vb == 0 ? a.lt - a.js : a.je - a.lb;
      } else {  // the lean rows
        j0 = a.lt + vb * a.H;
        h = min(a.H, a.lb + 1 - j0);
      }
    }
    (void)ns;
  }
  const int i0 = a.is + strip * FT_BX;
  StripCtx<T> sc;
  sc.j0 = j0;
  sc.h = h;
  sc.i = i0 - NG + lane;
  sc.ic = sc.i > a.ied ? a.ied : sc.i;
  sc.lo_q = sc.ic - a.isd;
  const int fi = sc.i < a.is ? a.is : (sc.i > a.ie + 1 ? a.ie + 1 : sc.i);
  sc.lo_fx = fi - a.is;
  sc.lo_mx = (sc.i > a.ie ? a.ie : (sc.i < a.is ? a.is : sc.i)) - a.is;
  sc.col_out = lane >= NG && lane < NG + FT_BX && sc.i <= a.ie;
  sc.lane_in = sc.i >= a.is && sc.i <= a.ie;
  sc.edge_strip = EK >= 2 && a.edge && (i0 - NG <= 2 || i0 - NG + FT_LANES - 1 >= a.npx - 2);
  if constexpr (sizeof(T) == 4) sc.ec = edge_coef_of(sc.edge_strip ? edge_role(sc.i, a.is, a.ie, a.npx, a.edge) : 0);
  sc.edge_rows = EK != 0 && EK <= 4;
  sc.gen_src = EK != 0;
  sc.corner_wg = a.corners != 0u && (strip == 0 || strip == a.nstrip - 1) &&
                 (EK == 2 ? (vb == 0 || vb == a.nvblk - 1) : (EK == 3 || EK == 4));
  sc.scale = nspl > 1;
  sc.frac = T(1) / T(nspl);
  sc.nix = a.ie - a.is + 2;
  sc.niy = a.ie - a.is + 1;
  // this block's tile (the argument block itself stays unmodified: a kernel
  // argument written to is copied to the private stack)
  const long long tile = blockIdx.z;
  sc.tg = tile * a.tgs2;
  sc.tdx = tile * a.tgdx;
  sc.tdy = tile * a.tgdy;
  sc.tgf = tile * a.tgf;
  sc.cxk = a.cx + tile * a.tcx + (long long)k * a.sx;
  sc.cyk = a.cy + tile * a.tcy + (long long)k * a.sy;
  sc.mfxk = a.mfx + tile * a.tmx + (long long)k * a.smx;
  sc.mfyk = a.mfy + tile * a.tmy + (long long)k * a.smy;
  sc.dp1k = a.dp1 + tile * a.tdp + (long long)k * a.s2;
  const T *in;
  T *out;
  pass_buffers(a, nspl, in, out);
  const int per_round = W * NT;
  const int rounds = (a.nq + per_round - 1) / per_round;
  const int nrow = sc.h + 2 * NG;
  for (int rd = 0; rd < rounds; rd++) {
    TrState<T> st[NT];
    bool act[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int iq = rd * per_round + wave * NT + n;
      act[n] = iq < a.nq;
      const long long qoff = (long long)(act[n] ? iq : 0) * a.npz * a.s2 + (long long)k * a.s2;
      const long long toff = (long long)blockIdx.z * a.tq;
      st[n].qsrc0 = in + qoff;
      st[n].qin = in + toff + qoff;
      st[n].qhalo = a.q_in + toff + qoff;
      st[n].qout = out + toff + qoff;
      if constexpr (sizeof(T) == 4) {
        st[n].qout_r = __builtin_amdgcn_make_buffer_rsrc(st[n].qout, (short)0, 0x7fffffff, 0x00020000);
        st[n].qin_r = __builtin_amdgcn_make_buffer_rsrc(const_cast<T *>(st[n].qin), (short)0, 0x7fffffff, 0x00020000);
      }
      st[n].fy2_prev = st[n].Fy_prev = T(0);
    }
    LvState<T> lv;
    lv.yfx_prev = lv.mfy_prev = T(0);
    const bool any = act[0];
    if constexpr (sizeof(T) == 4) {
      constexpr int QN = ek_qn(EK);
      static_assert(LR_AHEAD == 3 && LR_DEPTH >= LR_AHEAD + 5, "ring_wait covers up to two rows ahead; rows t-4..t+3 kept");
      __shared__ LdsRing<W, NT, QN> ring;
      // the previous round's last reads of every slot are done before the
      // prologue overwrites them
      if (rd > 0) ring_wait<W, NT, QN>(0);
      const LeanRows<W> lr = lean_rows<W>(a, sc, wave);
      auto issue = [&](int t) {
        if constexpr (EK == 0) {
          ring_issue_lean<W, NT>(a, sc, st, lr, ring, wave, t);
        } else if constexpr (EK >= 5) {  // lean rows: no row index clamps
          ring_issue_mixed<W, NT, EK>(a, sc, st, lr, ring, wave, t);
        } else {
          // rows whose row indices do not clamp take the affine addresses
          // for the level arrays; the q sources keep their per-lane rules
          const int r = sc.j0 - NG + t;
          if (r >= a.js + NG && r <= a.je)
            ring_issue_mixed<W, NT, EK>(a, sc, st, lr, ring, wave, t);
          else
            ring_issue<W, NT, EK>(a, sc, st, ring, wave, t);
        }
      };
      // The lean kinds (0, and 5 / 6 in the edge strips, whose blocks are H
      // rows) run whole groups of six rows: the rows past the block read ring
      // slots that were not reloaded (in-bounds LDS) and store nothing
      // (x.out_row is false there), and the loop carries no per-row guard.
      // Every wave of the workgroup steps through the same rows, so the
      // barriers in ring_wait pair up.
      constexpr bool PAD = EK == 0 || EK >= 5;
      const int nrun = PAD ? (nrow + 5) / 6 * 6 : nrow;
      for (int t = 0; t < LR_AHEAD && t < nrow; t++) issue(t);
#define FV3_RROW(U, TW)                                                              \
  if (PAD || t0 + U < nrow) {                                                        \
    const int t = t0 + U;                                                            \
    ring_wait<W, NT, QN>(min(LR_AHEAD - 1, nrow - 1 - t));                           \
    if (t + LR_AHEAD < nrow) issue(t + LR_AHEAD);                                    \
    RowIn<T, NT> in;                                                                 \
    ring_read<W, NT, QN>(ring, wave, lane, t, in);                                   \
    row_step<T, NT, U, HO, ek_side(EK), TW>(a, sc, lv, st, act, t, in);              \
  }
      if (any) {
        // the lean kind runs the block's first six input rows (the y-pipes'
        // warm-up) as their own unrolled group; the edge kinds would need
        // more VGPRs with it (kinds 5 and 6: 142 against 95)
        if constexpr (EK == 0) {
          constexpr int t0 = 0;
          FV3_RROW(0, 0) FV3_RROW(1, 1) FV3_RROW(2, 2) FV3_RROW(3, 3) FV3_RROW(4, 4) FV3_RROW(5, 5)
        }
        for (int t0 = EK == 0 ? 6 : 0; t0 < nrun; t0 += 6) {
          FV3_RROW(0, 6) FV3_RROW(1, 6) FV3_RROW(2, 6) FV3_RROW(3, 6) FV3_RROW(4, 6) FV3_RROW(5, 6)
        }
      } else {  // no tracer of this wave in this round: its loads and barriers only
        for (int t = 0; t < nrun; t++) {
          ring_wait<W, NT, QN>(min(LR_AHEAD - 1, nrow - 1 - t));
          if (t + LR_AHEAD < nrow) issue(t + LR_AHEAD);
        }
      }
#undef FV3_RROW
      continue;
    }
    // rows are fetched one ahead; the bare s_barrier keeps the waves of the
    // workgroup (same level data, L1 reuse) in step without waiting for the
    // loads in flight -- the waves share no memory
    // double buffer by row parity (6 rows per unrolled step: parity is U's)
    RowIn<T, NT> buf[2];
    if (any) row_fetch<T, NT>(a, sc, st, act, 0, buf[0]);
    for (int t0 = 0; t0 < nrow; t0 += 6) {
#define FV3_ROW(U)                                                                   \
  if (any && t0 + U < nrow) {                                                        \
    if (t0 + U + 1 < nrow) row_fetch<T, NT>(a, sc, st, act, t0 + U + 1, buf[(U + 1) & 1]); \
    row_step<T, NT, U, HO, ek_side(EK)>(a, sc, lv, st, act, t0 + U, buf[U & 1]); \
  }                                                                                  \
  if (W > 1) __builtin_amdgcn_s_barrier();
      FV3_ROW(0) FV3_ROW(1) FV3_ROW(2) FV3_ROW(3) FV3_ROW(4) FV3_ROW(5)
#undef FV3_ROW
    }
  }
}

}  // namespace fv3
