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
  int i, ic, j0, nix, niy, lo_q, lo_fx, lo_mx;  // lane column; lane offsets in the row layouts
  bool col_out, lane_in, edge_strip, corner_wg, scale;
  T frac;
  const T *cxk, *cyk, *mfxk, *mfyk;
  const double *dp1k;
};

// per tracer streaming state
template <typename T>
struct TrState {
  YPipe<T> py, pj;
  T fx2_r[6];
  T fy2_prev, Fy_prev;
  const T *qin, *qhalo;
  T *qout;
};

// level data carried between rows
template <typename T>
struct LvState {
  T area_r[6], cx_r[6];
  T yfx_prev, mfy_prev;
};

// One input row r = j0 - 3 + t of the streaming pass; U = t mod 6.
template <typename T, int NT, int U>
__device__ __forceinline__ void row_step(const FusedArgs<T> &a, const StripCtx<T> &sc,
                                         LvState<T> &lv, TrState<T> (&st)[NT], const bool (&act)[NT],
                                         int t) {
  constexpr bool fp64 = sizeof(T) == 8;
  constexpr int M3 = (U + 3) % 6;  // slot of row r-3
  const int nid = a.nid;
  const int r = sc.j0 - NG + t;
  const int rc = r > a.jed ? a.jed : r;
  const int m = r - 3;
  const int mc = m < a.js ? a.js : (m > a.je ? a.je : m);
  const int f = r - 2;
  const int fc = f < a.js ? a.js : (f > a.je + 1 ? a.je + 1 : f);
  const long long orow = (long long)(rc - a.jsd) * nid;  // row rc of (isd:ied, jsd:jed)
  const bool row_in = rc >= a.js && rc <= a.je;
  const int ord_in = a.hord == 10 ? 8 : a.hord;

  // ---- level data of this row (shared by the NT tracers)
  const T area_r = a.area[orow + sc.lo_q];
  lv.area_r[U] = area_r;
  const T cyu = sc.cyk[(long long)(fc - a.js) * nid + sc.lo_q];
  const T cys = sc.scale ? cyu * sc.frac : cyu;
  T yfx;
  {
    const bool pos = cyu > T(0);
    const long long fo = (long long)(fc - a.jsd) * nid + sc.lo_q;
    if constexpr (fp64) {
      const long long uo = pos ? fo - nid : fo;
      yfx = cyu * a.dya[uo] * a.dx[fo] * a.sin_sg[(pos ? 3 : 1) * a.ms2 + uo];
    } else {
      yfx = cyu * (pos ? a.gyp[fo] : a.gym[fo]);
    }
    if (sc.scale) yfx = yfx * sc.frac;
  }
  const T cxr_u = sc.cxk[(long long)(rc - a.jsd) * sc.nix + sc.lo_fx];
  lv.cx_r[U] = cxr_u;
  const T cxr = sc.scale ? cxr_u * sc.frac : cxr_u;
  T xfx;
  {
    const bool pos = cxr_u > T(0);
    if constexpr (fp64) {
      const long long uo = orow + sc.lo_fx + (a.is - a.isd) - (pos ? 1 : 0);
      xfx = cxr_u * a.dxa[uo] * a.dy[(long long)(rc - a.jsd) * (nid + 1) + sc.lo_fx + (a.is - a.isd)] *
            a.sin_sg[(pos ? 2 : 0) * a.ms2 + uo];
    } else {
      const long long fo = (long long)(rc - a.jsd) * (nid + 1) + sc.lo_fx + (a.is - a.isd);
      xfx = cxr_u * (pos ? a.gxp[fo] : a.gxm[fo]);
    }
    if (sc.scale) xfx = xfx * sc.frac;
  }
  const T xfxp = from_right(xfx);
  const T rax = area_r + xfx - xfxp;
  const T area_m = lv.area_r[M3];
  const T ray = area_m + lv.yfx_prev - yfx;
  T rrax = T(0), rray = T(0);
  if constexpr (!fp64) {
    rrax = __builtin_amdgcn_rcpf(rax);
    rray = __builtin_amdgcn_rcpf(ray);
  }
  const T cxm_u = lv.cx_r[M3];  // cx of row m (row r-3 was loaded 3 steps ago)
  const T cxm = sc.scale ? cxm_u * sc.frac : cxm_u;
  const T mfy_f = sc.mfyk[(long long)(fc - a.js) * sc.niy + sc.lo_mx];
  const T mfys = sc.scale ? mfy_f * sc.frac : mfy_f;
  const T mfx_m = sc.mfxk[(long long)(mc - a.js) * sc.nix + sc.lo_fx];
  const T mfxs = sc.scale ? mfx_m * sc.frac : mfx_m;
  const T mfxp = from_right(mfxs);
  T dxa_m = T(1), dxa_r = T(1);
  if (sc.edge_strip) {
    dxa_m = a.dxa[(long long)(mc - a.jsd) * nid + sc.lo_q];
    dxa_r = a.dxa[orow + sc.lo_q];
  }
  // cube-edge role of cell r-2 along the columns (uniform across the wave)
  const int role_y = edge_role(r - 2, a.js, a.je, a.npy, a.edge);
  T dd0 = T(0), dd1 = T(0), dd2 = T(0), dd3 = T(0), dd4 = T(0);
  if (role_y != 0) {
    auto dya_at = [&](int jj) {
      jj = jj < a.jsd ? a.jsd : (jj > a.jed ? a.jed : jj);
      return a.dya[(long long)(jj - a.jsd) * nid + sc.lo_q];
    };
    dd0 = dya_at(r - 4);
    dd1 = dya_at(r - 3);
    dd2 = dya_at(r - 2);
    dd3 = dya_at(r - 1);
    dd4 = dya_at(r);
  }
  // output row: thickness and area factors shared by the tracers
  const bool out_row = t >= 2 * NG && m >= sc.j0 && m < sc.j0 + a.H && m <= a.je;
  T rarea = T(0), d1 = T(0), rdp2 = T(0), dp2 = T(0);
  long long co = 0;
  if (out_row) {
    co = (long long)(mc - a.jsd) * nid + sc.lo_q;
    rarea = a.rarea[co];
    const T ddiv = ((mfxs - mfxp) + (lv.mfy_prev - mfys)) * rarea;
    T x = (T)sc.dp1k[co];
    for (int s = 1; s < a.it; s++) x = x + ddiv;
    d1 = x;
    dp2 = x + ddiv;
    if constexpr (!fp64) rdp2 = __builtin_amdgcn_rcpf(dp2);
  }

  // ---- the tracers
#pragma unroll
  for (int n = 0; n < NT; n++) {
    if (!act[n]) continue;
    TrState<T> &s = st[n];
    const long long o = orow + sc.lo_q;
    T qy = (sc.lane_in && row_in) ? s.qin[o] : s.qhalo[o];
    T qx = qy;
    if (sc.corner_wg && !row_in && !sc.lane_in) {
      int si, sj;
      if (corner_src(sc.ic, rc, a.npx, a.npy, 2, a.corners, si, sj)) {
        qy = s.qhalo[(long long)(sj - a.jsd) * nid + (si - a.isd)];
        corner_src(sc.ic, rc, a.npx, a.npy, 1, a.corners, si, sj);
        qx = s.qhalo[(long long)(sj - a.jsd) * nid + (si - a.isd)];
      }
    }
    // inner y-sweep at face r-2 and q_i of row r-3
    const T fy2 = ypipe_step<T, U>(s.py, qy, cys, ord_in, role_y, dd0, dd1, dd2, dd3, dd4);
    const T qm = s.py.w[M3];
    T q_i;
    if constexpr (fp64)
      q_i = (qm * area_m + lv.yfx_prev * s.fy2_prev - yfx * fy2) / ray;
    else
      q_i = (qm * area_m + lv.yfx_prev * s.fy2_prev - yfx * fy2) * rray;
    // outer x-sweep of row r-3, inner x-sweep of row r, q_j of row r
    const T fx = xsweep(q_i, cxm, sc.i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, a.hord, dxa_m);
    const T fx2 = xsweep(qx, cxr, sc.i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, ord_in, dxa_r);
    s.fx2_r[U] = fx2;
    const T fx2p = from_right(fx2);
    T q_j;
    if constexpr (fp64)
      q_j = (qx * area_r + xfx * fx2 - xfxp * fx2p) / rax;
    else
      q_j = (qx * area_r + xfx * fx2 - xfxp * fx2p) * rrax;
    // outer y-sweep at face r-2, fluxes, update of row r-3
    const T fy = ypipe_step<T, U>(s.pj, q_j, cys, a.hord, role_y, dd0, dd1, dd2, dd3, dd4);
    const T Fy = T(0.5) * (fy + fy2) * mfys;
    const T Fx = T(0.5) * (fx + s.fx2_r[M3]) * mfxs;
    const T Fxp = from_right(Fx);
    if (out_row && sc.col_out) {
      if constexpr (fp64)
        s.qout[co] = (qm * d1 + (Fx - Fxp + s.Fy_prev - Fy) * rarea) / dp2;
      else
        s.qout[co] = (qm * d1 + (Fx - Fxp + s.Fy_prev - Fy) * rarea) * rdp2;
    }
    s.fy2_prev = fy2;
    s.Fy_prev = Fy;
  }
  lv.yfx_prev = yfx;
  lv.mfy_prev = mfys;
}

template <typename T, int W, int NT>
__global__ void __launch_bounds__(FT_LANES *W) k_fused_tp(FusedArgs<T> a) {
  const int lane = threadIdx.x & (FT_LANES - 1), wave = threadIdx.x / FT_LANES;
  const int k = blockIdx.y;
  const int nspl = a.nsplit[k];
  // whole block: this level is done, or needs more passes than are launched
  // (reported by the caller's nsplit check; nothing is written for it)
  if (a.it > nspl || nspl > a.cap) return;
  const int strip = blockIdx.x % a.nstrip, vb = blockIdx.x / a.nstrip;
  const int i0 = a.is + strip * FT_BX;
  StripCtx<T> sc;
  sc.j0 = a.js + vb * a.H;
  sc.i = i0 - NG + lane;
  sc.ic = sc.i > a.ied ? a.ied : sc.i;
  sc.lo_q = sc.ic - a.isd;
  const int fi = sc.i < a.is ? a.is : (sc.i > a.ie + 1 ? a.ie + 1 : sc.i);
  sc.lo_fx = fi - a.is;
  sc.lo_mx = (sc.i > a.ie ? a.ie : (sc.i < a.is ? a.is : sc.i)) - a.is;
  sc.col_out = lane >= NG && lane < NG + FT_BX && sc.i <= a.ie;
  sc.lane_in = sc.i >= a.is && sc.i <= a.ie;
  sc.edge_strip = a.edge && (i0 - NG <= 2 || i0 - NG + FT_LANES - 1 >= a.npx - 2);
  sc.corner_wg = a.corners != 0u && (strip == 0 || strip == a.nstrip - 1) && (vb == 0 || vb == a.nvblk - 1);
  sc.scale = nspl > 1;
  sc.frac = T(1) / T(nspl);
  sc.nix = a.ie - a.is + 2;
  sc.niy = a.ie - a.is + 1;
  sc.cxk = a.cx + (long long)k * a.sx;
  sc.cyk = a.cy + (long long)k * a.sy;
  sc.mfxk = a.mfx + (long long)k * a.smx;
  sc.mfyk = a.mfy + (long long)k * a.smy;
  sc.dp1k = a.dp1 + (long long)k * a.s2;
  const T *in;
  T *out;
  pass_buffers(a, nspl, in, out);
  const int per_round = W * NT;
  const int rounds = (a.nq + per_round - 1) / per_round;
  const int nrow = a.H + 2 * NG;
  for (int rd = 0; rd < rounds; rd++) {
    TrState<T> st[NT];
    bool act[NT];
#pragma unroll
    for (int n = 0; n < NT; n++) {
      const int iq = rd * per_round + wave * NT + n;
      act[n] = iq < a.nq;
      const long long qoff = (long long)(act[n] ? iq : 0) * a.npz * a.s2 + (long long)k * a.s2;
      st[n].qin = in + qoff;
      st[n].qhalo = a.q_in + qoff;
      st[n].qout = out + qoff;
      st[n].fy2_prev = st[n].Fy_prev = T(0);
    }
    LvState<T> lv;
    lv.yfx_prev = lv.mfy_prev = T(0);
    const bool any = act[0];
    for (int t0 = 0; t0 < nrow; t0 += 6) {
#define FV3_ROW(U)                                                          \
  if (any && t0 + U < nrow) row_step<T, NT, U>(a, sc, lv, st, act, t0 + U); \
  if (W > 1) __syncthreads();
      FV3_ROW(0) FV3_ROW(1) FV3_ROW(2) FV3_ROW(3) FV3_ROW(4) FV3_ROW(5)
#undef FV3_ROW
    }
  }
}

}  // namespace fv3
