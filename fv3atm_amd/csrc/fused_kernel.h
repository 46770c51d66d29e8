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

// per (strip, block, level) constants of a workgroup
template <typename T>
struct StripCtx {
  int i, ic, j0, nix, niy;
  bool col_out, edge_strip, scale;
  T frac;
  const T *cxk, *cyk, *mfxk, *mfyk, *dyac;
  const double *dp1k;
};

// per tracer streaming state
template <typename T>
struct TrState {
  YPipe<T> py, pj;
  T area_r[6], fx2_r[6];
  T yfx_prev, fy2_prev, Fy_prev, mfy_prev;
  const T *qin, *qhalo;
  T *qout;
};

// One input row r = j0 - 3 + t of the streaming pass; U = t mod 6.
template <typename T, int U>
__device__ __forceinline__ void row_step(const FusedArgs<T> &a, const StripCtx<T> &sc,
                                         TrState<T> &st, int t) {
  constexpr bool fp64 = sizeof(T) == 8;
  constexpr int M3 = (U + 3) % 6;  // slot of row r-3
  const int nid = a.nid;
  const int i = sc.i, ic = sc.ic;
  const int r = sc.j0 - NG + t;
  const int rc = r > a.jed ? a.jed : r;
  // ---- q of row r, y-sweep and x-sweep variants (they differ in the corner blocks)
  T qy, qx;
  {
    const bool interior = i >= a.is && i <= a.ie && rc >= a.js && rc <= a.je;
    int si, sj;
    const long long o = (long long)(rc - a.jsd) * nid + (ic - a.isd);
    if (interior) {
      qy = qx = st.qin[o];
    } else if (corner_src(ic, rc, a.npx, a.npy, 2, a.corners, si, sj)) {
      qy = st.qhalo[(long long)(sj - a.jsd) * nid + (si - a.isd)];
      corner_src(ic, rc, a.npx, a.npy, 1, a.corners, si, sj);
      qx = st.qhalo[(long long)(sj - a.jsd) * nid + (si - a.isd)];
    } else {
      qy = qx = st.qhalo[o];
    }
  }
  const long long mo = (long long)(rc - a.jsd) * nid + (ic - a.isd);
  st.area_r[U] = a.area[mo];
  // ---- inner y-sweep: fy2 at face f = r-2 of this column
  const int f = r - 2;
  const int fc = f < a.js ? a.js : (f > a.je + 1 ? a.je + 1 : f);
  const T cyu = sc.cyk[(long long)(fc - a.js) * nid + (ic - a.isd)];
  const T cys = sc.scale ? cyu * sc.frac : cyu;
  T yfx;
  {
    const bool pos = cyu > T(0);
    const long long fo = (long long)(fc - a.jsd) * nid + (ic - a.isd);
    if (fp64) {
      const long long uo = pos ? fo - nid : fo;
      yfx = cyu * a.dya[uo] * a.dx[fo] * a.sin_sg[(pos ? 3 : 1) * a.ms2 + uo];
    } else {
      yfx = cyu * (pos ? a.gyp : a.gym)[fo];
    }
    if (sc.scale) yfx = yfx * sc.frac;
  }
  const int ord_in = a.hord == 10 ? 8 : a.hord;
  const T fy2 = ypipe_step<T, U>(st.py, qy, r, cys, a.js, a.je, a.npy, a.edge, ord_in, sc.dyac, nid,
                                 a.jsd, a.jed);
  // ---- q_i of row m = r-3 and the outer x-sweep fx
  const int m = r - 3;
  const T area_m = st.area_r[M3];
  const T ray = area_m + st.yfx_prev - yfx;
  const T qm = st.py.w[M3];
  const T q_i = (qm * area_m + st.yfx_prev * st.fy2_prev - yfx * fy2) / ray;
  const int mc = m < a.js ? a.js : (m > a.je ? a.je : m);
  const int fi = i < a.is ? a.is : (i > a.ie + 1 ? a.ie + 1 : i);
  const T cxm_u = sc.cxk[(long long)(mc - a.jsd) * sc.nix + (fi - a.is)];
  const T cxm = sc.scale ? cxm_u * sc.frac : cxm_u;
  const T dxa_m = a.dxa[(long long)(mc - a.jsd) * nid + (ic - a.isd)];
  const T fx = xsweep(q_i, cxm, i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, a.hord, dxa_m);
  // ---- inner x-sweep fx2 of row r and q_j of row r
  const T cxr_u = sc.cxk[(long long)(rc - a.jsd) * sc.nix + (fi - a.is)];
  const T cxr = sc.scale ? cxr_u * sc.frac : cxr_u;
  T xfx;
  {
    const bool pos = cxr_u > T(0);
    const long long fo = (long long)(rc - a.jsd) * (nid + 1) + (fi - a.isd);
    if (fp64) {
      const long long uo = (long long)(rc - a.jsd) * nid + ((pos ? fi - 1 : fi) - a.isd);
      xfx = cxr_u * a.dxa[uo] * a.dy[fo] * a.sin_sg[(pos ? 2 : 0) * a.ms2 + uo];
    } else {
      xfx = cxr_u * (pos ? a.gxp : a.gxm)[fo];
    }
    if (sc.scale) xfx = xfx * sc.frac;
  }
  const T dxa_r = a.dxa[mo];
  const T fx2 = xsweep(qx, cxr, i, a.is, a.ie, a.npx, a.edge, sc.edge_strip, ord_in, dxa_r);
  st.fx2_r[U] = fx2;
  const T fx2p = shdn(fx2, 1), xfxp = shdn(xfx, 1);
  const T rax = st.area_r[U] + xfx - xfxp;
  const T q_j = (qx * st.area_r[U] + xfx * fx2 - xfxp * fx2p) / rax;
  // ---- outer y-sweep fy at face r-2 on q_j
  const T fy = ypipe_step<T, U>(st.pj, q_j, r, cys, a.js, a.je, a.npy, a.edge, a.hord, sc.dyac, nid,
                                a.jsd, a.jed);
  const int ii = i > a.ie ? a.ie : (i < a.is ? a.is : i);
  const T mfy_f = sc.mfyk[(long long)(fc - a.js) * sc.niy + (ii - a.is)];
  const T mfys = sc.scale ? mfy_f * sc.frac : mfy_f;
  const T Fy = T(0.5) * (fy + fy2) * mfys;
  // ---- x flux of row m and the flux-form update of row m
  const T mfx_m = sc.mfxk[(long long)(mc - a.js) * sc.nix + (fi - a.is)];
  const T mfxs = sc.scale ? mfx_m * sc.frac : mfx_m;
  const T Fx = T(0.5) * (fx + st.fx2_r[M3]) * mfxs;
  const T Fxp = shdn(Fx, 1), mfxp = shdn(mfxs, 1);
  if (t >= 2 * NG && sc.col_out && m >= sc.j0 && m < sc.j0 + a.H && m <= a.je) {
    const long long co = (long long)(m - a.jsd) * nid + (i - a.isd);
    const T rarea = a.rarea[co];
    if (fp64) {
      const T dd = ((mfxs - mfxp) + (st.mfy_prev - mfys)) * rarea;
      T d1 = (T)sc.dp1k[co];
      for (int s = 1; s < a.it; s++) d1 = d1 + dd;
      const T dp2 = d1 + dd;
      st.qout[co] = (qm * d1 + (Fx - Fxp + st.Fy_prev - Fy) * rarea) / dp2;
    } else {
      const double dd = (((double)mfxs - (double)mfxp) + ((double)st.mfy_prev - (double)mfys)) * (double)rarea;
      double d1 = sc.dp1k[co];
      for (int s = 1; s < a.it; s++) d1 = d1 + dd;
      const double dp2 = d1 + dd;
      st.qout[co] = (T)(((double)qm * d1 + (double)((Fx - Fxp + st.Fy_prev - Fy) * rarea)) / dp2);
    }
  }
  st.yfx_prev = yfx;
  st.fy2_prev = fy2;
  st.Fy_prev = Fy;
  st.mfy_prev = mfys;
}

template <typename T, int W>
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
  sc.col_out = lane >= NG && lane < NG + FT_BX && sc.i <= a.ie;
  sc.edge_strip = a.edge && (i0 - NG <= 2 || i0 - NG + FT_LANES - 1 >= a.npx - 2);
  sc.scale = nspl > 1;
  sc.frac = T(1) / T(nspl);
  sc.nix = a.ie - a.is + 2;
  sc.niy = a.ie - a.is + 1;
  sc.cxk = a.cx + (long long)k * a.sx;
  sc.cyk = a.cy + (long long)k * a.sy;
  sc.mfxk = a.mfx + (long long)k * a.smx;
  sc.mfyk = a.mfy + (long long)k * a.smy;
  sc.dp1k = a.dp1 + (long long)k * a.s2;
  sc.dyac = a.dya + (sc.ic - a.isd);
  const T *in;
  T *out;
  pass_buffers(a, nspl, in, out);
  const int rounds = (a.nq + W - 1) / W;
  const int nrow = a.H + 2 * NG;
  for (int rd = 0; rd < rounds; rd++) {
    const int iq = wave + W * rd;
    const bool active = iq < a.nq;
    const long long qoff = (long long)(active ? iq : 0) * a.npz * a.s2 + (long long)k * a.s2;
    TrState<T> st;
    st.qin = in + qoff;
    st.qhalo = a.q_in + qoff;
    st.qout = out + qoff;
    st.yfx_prev = st.fy2_prev = st.Fy_prev = st.mfy_prev = T(0);
    for (int t0 = 0; t0 < nrow; t0 += 6) {
#define FV3_ROW(U)                                                 \
  if (active && t0 + U < nrow) row_step<T, U>(a, sc, st, t0 + U); \
  if (W > 1) __syncthreads();
      FV3_ROW(0) FV3_ROW(1) FV3_ROW(2) FV3_ROW(3) FV3_ROW(4) FV3_ROW(5)
#undef FV3_ROW
    }
  }
}

}  // namespace fv3
