// mapz_fast.hip -- FP32 vertical remap of the Lagrangian_to_Eulerian tracer
// subset (upstream fv_mapz.F90: mapn_tracer / map1_q2 for the tracers with
// scalar_profile iv=0, map_scalar iv=1 for pt), laid out for the fused step.
//
// Everything that depends only on the column's pressure thicknesses -- the
// Lagrangian and Eulerian interface pressures, the coefficients of the
// interface-value tridiagonal solve of cs_profile/scalar_profile, and for
// each Eulerian interface the Lagrangian layer that contains it -- is the
// same for every tracer of the column.  k_remap_prep computes it once per
// column into [level][column] arrays; k_remap_fields then runs one thread per
// (column, field) with 64 columns x 4 fields per workgroup, so the four
// waves of a workgroup read the same column data.
//
// Per field the solve is the upstream Thomas recurrence (elimination
// downward, substitution upward).  The downward pass keeps every B-th
// eliminated value in LDS; the upward pass regenerates each block of B
// levels in registers from its checkpoint, back-substitutes, applies the
// interface limits, the extremum flags and the kord 8/9/10 cell limiters
// through a rolling window of a few levels, and integrates the Eulerian
// layers bottom-up (the upstream sums top-down; the pieces are the same).
// Results match the FP64 form to FP32 rounding, not bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "fv3_internal.h"
#include "mapz_column.h"
#include "mapz_fast.h"

namespace fv3 {

namespace {

constexpr int RF_B = 8;       // levels per regenerated block
constexpr int RF_COLS = 64;   // columns per workgroup (one per lane)
constexpr int RF_WAVES = 4;   // fields per workgroup (one per wave)

struct PrepArgs {
  int km, ni, is, js, isd, jsd, nid;
  long long ncol, s2;
  double ptop;
  const double *dp1, *ak, *bk;
  double *delp_out;
  RemapCols cols;
};

// One thread per interior column.
__global__ void __launch_bounds__(256) k_remap_prep(const PrepArgs a) {
  const long long col = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (col >= a.ncol) return;
  const int ii = (int)(col % a.ni), jj = (int)(col / a.ni);
  const long long cell = (long long)(jj + a.js - a.jsd) * a.nid + (ii + a.is - a.isd);
  const double *dl = a.dp1 + cell;
  const long long s2 = a.s2, nc = a.ncol;
  const int km = a.km;
  const RemapCols &o = a.cols;
  // Lagrangian interfaces by running sums from ptop, thicknesses as their
  // differences (the Lagrangian_to_Eulerian pe1 / dp1 of the upstream)
  double p_k = a.ptop;             // PE1(1)
  double p_k1 = p_k + dl[0];       // PE1(2)
  double d_prev = p_k1 - p_k;      // D1(1)
  const double p_k2 = p_k1 + dl[s2];
  const double grat = (p_k2 - p_k1) / d_prev;
  const double bet = grat * (grat + 0.5);
  double g = (1.0 + grat * (grat + 1.5)) / bet;
  o.d4[col] = (float)((grat + grat) * (grat + 1.0));
  o.rbt[col] = (float)(1.0 / bet);
  o.gg[col] = (float)g;
  o.dp1[col] = (float)d_prev;
  double d4 = 0.0;
  p_k = p_k1;
  for (int k = 2; k <= km; k++) {
    const double p_n = p_k + dl[(long long)(k - 1) * s2];
    const double dk = p_n - p_k;
    d4 = d_prev / dk;
    const double bt = 2.0 + d4 + d4 - g;
    g = d4 / bt;
    const long long o_k = (long long)(k - 1) * nc + col;
    o.d4[o_k] = (float)d4;
    o.rbt[o_k] = (float)(1.0 / bt);
    o.gg[o_k] = (float)g;
    o.dp1[o_k] = (float)dk;
    d_prev = dk;
    p_k = p_n;
  }
  const double ps = p_k;  // PE1(km+1)
  const double a_bot = 1.0 + d4 * (d4 + 1.5);
  o.bot[col] = (float)(2.0 * d4 * (d4 + 1.0));
  o.bot[nc + col] = (float)a_bot;
  o.bot[2 * nc + col] = (float)(1.0 / (d4 * (d4 + 0.5) - a_bot * g));

  // Eulerian interfaces, target thickness, and the containing source layer
  // of every target interface (the first layer l with pe2 <= PE1(l+1), the
  // choice the upstream search makes when an interface falls on a source
  // interface)
  auto pe2_of = [&](int k) -> double {
    if (k == 1) return a.ptop;
    if (k == km + 1) return ps;
    return a.ak[k - 1] + a.bk[k - 1] * ps;
  };
  int l = 1;
  double lo = a.ptop, hi = a.ptop + dl[0];  // PE1(l), PE1(l+1)
  double pe2_k = a.ptop;
  float s_prev = 0.f, wtop_prev = 0.f, rdp2_prev = 0.f;
  int l_prev = 1;
  for (int k = 1; k <= km + 1; k++) {
    const double p = pe2_k;
    while (l < km && p > hi) {
      l++;
      lo = hi;
      hi = lo + dl[(long long)(l - 1) * s2];
    }
    if (k >= 2) {  // close target k-1: its bottom piece lies in layer l
      const long long o_t = (long long)(k - 2) * nc + col;
      o.cut[o_t] = make_float4(s_prev, wtop_prev, (float)(p - lo), rdp2_prev);
      o.lk[o_t] = l_prev;
    }
    if (k <= km) {
      const double p_next = pe2_of(k + 1);
      const double dp2 = p_next - p;
      a.delp_out[(long long)(k - 1) * s2 + cell] = dp2;
      s_prev = (float)((p - lo) / (hi - lo));
      wtop_prev = (float)(hi - p);
      rdp2_prev = (float)(1.0 / dp2);
      l_prev = l;
      pe2_k = p_next;
    }
  }
}

struct FieldArgs {
  int km, nb, ni, is, js, isd, jsd, nid, nq, nf, nfg, kord_tr, kord_tm;
  unsigned nwg;
  long long ncol, s2;
  float t_min;
  const float *q_src, *pt_src;
  float *q_dst, *pt_dst;
  RemapCols cols;
  // byte offsets of rbt gg dp1 bot lk from d4 (one carved block), and the
  // byte sizes of that block (d4 .. lk) and of cut
  unsigned o_rbt, o_gg, o_dp1, o_bot, o_lk, b_cols, b_cut;
};

// Raw buffer access: a wave-uniform descriptor per array, the lane's byte
// offset in a VGPR and the level's byte offset in an SGPR, so that no load
// needs per-lane 64-bit address arithmetic.  (Every array of a launch is
// below 2 GiB; descriptor word 3 is the gfx9 raw 32-bit data format.)
using Rsrc = __amdgpu_buffer_rsrc_t;

__device__ __forceinline__ Rsrc make_rsrc(const void *p, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float ldf(Rsrc r, unsigned vo, unsigned so) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ int ldi(Rsrc r, unsigned vo, unsigned so) {
  return (int)__builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void stf(Rsrc r, unsigned vo, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, 0, 0);
}

// the column arrays of one launch: level k's value of the lane's column is
// at lane offset vc (= 4 col) and scalar offset o_X + (k-1) lvl
struct ColBuf {
  Rsrc r, rc;  // d4 .. lk; cut
  unsigned vc, lvl, o_rbt, o_gg, o_dp1, o_bot, o_lk;
  __device__ __forceinline__ float d4(int k) const { return ldf(r, vc, (unsigned)(k - 1) * lvl); }
  __device__ __forceinline__ float rbt(int k) const { return ldf(r, vc, o_rbt + (unsigned)(k - 1) * lvl); }
  __device__ __forceinline__ float gg(int k) const { return ldf(r, vc, o_gg + (unsigned)(k - 1) * lvl); }
  __device__ __forceinline__ float dp1(int k) const { return ldf(r, vc, o_dp1 + (unsigned)(k - 1) * lvl); }
  __device__ __forceinline__ float bot(int m) const { return ldf(r, vc, o_bot + (unsigned)m * lvl); }
};

__device__ __forceinline__ void huynh2f(float a1, float &a2, float &a3, float gk, float gk1, float gk2,
                                        float gkm1) {
  huynh2<float>(a1, a2, a3, gk, gk1, gk2, gkm1);
}

// mean of the cell parabola q(s) = c0 + s (a3 - a2 + a4) - a4 s^2 over [sa, sb]
__device__ __forceinline__ float pmean(float c0, float c1, float c2, float sa, float sb) {
  const float ss = sa + sb;
  return c0 + c1 * ss - c2 * (sb * ss + sa * sa);
}

struct TargetRow {
  float4 d;  // s (top interface fraction), wtop, wbot, 1/dp2
  int l;     // source layer of the top interface
};

// target row at the lane's level offset tvo (= (kk-1) lvl + vc); the four
// dword loads at consecutive offsets compile to one dwordx4 buffer load
__device__ __forceinline__ TargetRow load_target(const ColBuf &cb, unsigned tvo) {
  TargetRow t;
  const unsigned v = tvo * 4u;
  t.d = make_float4(ldf(cb.rc, v, 0), ldf(cb.rc, v + 4u, 0), ldf(cb.rc, v + 8u, 0), ldf(cb.rc, v + 12u, 0));
  t.l = ldi(cb.r, tvo, cb.o_lk);
  return t;
}

// Limited cell coefficients of an interior source layer (3 <= c <= km-2).
// e_lo/e_c/e_hi: flag bits (1 extm, 2 ext5, 4 ext6) of cells c-1, c, c+1.
__device__ __forceinline__ void cell_limit_interior(int iv, int akord, float qmin, float &a1, float &a2, float &a3,
                                                    float &a4, int e_lo, int e_c, int e_hi, float gkm1, float gk,
                                                    float gk1, float gk2) {
  if (akord < 9) {
    huynh2f(a1, a2, a3, gk, gk1, gk2, gkm1);
    a4 = 3.f * (2.f * a1 - (a2 + a3));
  } else if (akord == 9) {
    const bool em = (e_c & 1) != 0;
    if ((em && (e_lo & 1)) || (em && (e_hi & 1)) || (em && a1 < qmin)) {
      a2 = a1;
      a3 = a1;
      a4 = 0.f;
    } else {
      a4 = 6.f * a1 - 3.f * (a2 + a3);
      if (fabsf(a4) > fabsf(a2 - a3)) {
        huynh2f(a1, a2, a3, gk, gk1, gk2, gkm1);
        a4 = 6.f * a1 - 3.f * (a2 + a3);
      }
    }
  } else {
    if (e_c & 2) {
      if ((e_lo & 2) || (e_hi & 2) || a1 < qmin) {
        a2 = a1;
        a3 = a1;
      } else if ((e_lo & 4) || (e_hi & 4)) {
        huynh2f(a1, a2, a3, gk, gk1, gk2, gkm1);
      }
    } else if (e_c & 4) {
      if ((e_lo & 2) || (e_hi & 2)) huynh2f(a1, a2, a3, gk, gk1, gk2, gkm1);
    }
    a4 = 3.f * (2.f * a1 - (a2 + a3));
  }
  if (iv == 0) cs_lim1<float>(a1, a2, a3, a4, (e_c & 1) != 0, 0);
}

// Limited cell coefficients of source layer c (2 <= c <= km) from the window.
__device__ __forceinline__ void cell_limit(int c, int km, int iv, int akord, float qmin, float &a1,
                                           float &a2, float &a3, float &a4, int e_lo, int e_c,
                                           int e_hi, float gkm1, float gk, float gk1, float gk2) {
  if (akord > 16) {
    a4 = 3.f * (2.f * a1 - (a2 + a3));
    return;
  }
  if (c == km) {
    if (iv == 0) a3 = fmaxf(0.f, a3);
    a4 = 3.f * (2.f * a1 - (a2 + a3));
    cs_lim1<float>(a1, a2, a3, a4, (e_c & 1) != 0, 1);
  } else if (c == km - 1 || c == 2) {
    a4 = 3.f * (2.f * a1 - (a2 + a3));
    cs_lim1<float>(a1, a2, a3, a4, (e_c & 1) != 0, 2);
  } else {
    cell_limit_interior(iv, akord, qmin, a1, a2, a3, a4, e_lo, e_c, e_hi, gkm1, gk, gk1, gk2);
  }
}

// Mapping state of one (column, field): the open target kk (whose bottom
// interface has been passed), its partial sum, the row of target kk, and
// the lane's byte offsets of target kk in the column arrays (tvo) and in the
// destination (dvo).  The row of the next target is loaded as soon as kk
// closes; it is first needed a whole source layer later, which hides the load.
struct MapState {
  int kk;
  float acc;
  TargetRow cur;
  unsigned tvo, dvo;
};

// target kk's top interface lies in the current source layer: its mean is
// complete (the layer's piece above the bottom edge sb, plus what it already
// held); store it and move to target kk-1
__device__ __forceinline__ void close_target(MapState &m, float c0, float c1, float c2, bool &bot_in,
                                             float &sb, const ColBuf &cb, Rsrc dst, unsigned s2b,
                                             bool act) {
  const float s = m.cur.d.x;
  float v;
  if (bot_in)
    v = pmean(c0, c1, c2, s, sb);
  else
    v = (m.acc + m.cur.d.y * pmean(c0, c1, c2, s, 1.f)) * m.cur.d.w;
  if (act) stf(dst, m.dvo, v);
  m.acc = 0.f;
  bot_in = true;
  sb = s;
  m.kk--;
  m.tvo -= cb.lvl;
  m.dvo -= s2b;
  if (m.kk >= 1) m.cur = load_target(cb, m.tvo);
}

__device__ __forceinline__ void map_layer(MapState &m, int c, int km, float a1, float a2, float a3,
                                          float a4, float dp1c, const ColBuf &cb, Rsrc dst, unsigned s2b,
                                          bool act) {
  const float c0 = a2, c1 = 0.5f * (a4 + a3 - a2), c2 = a4 * (1.0f / 3.0f);
  bool bot_in = c == km;  // the bottom interface km+1 sits at s = 1 of layer km
  float sb = 1.f;
  // a source layer holds the top interfaces of zero, one or (rarely) more
  // targets: the first two closings as nested conditionals, the rest in a
  // loop that only a column with a third one enters (a loop around every
  // closing costs register copies of the whole state at each entry and exit)
  if (m.kk >= 1 && m.cur.l == c) {
    close_target(m, c0, c1, c2, bot_in, sb, cb, dst, s2b, act);
    if (m.kk >= 1 && m.cur.l == c) {
      close_target(m, c0, c1, c2, bot_in, sb, cb, dst, s2b, act);
      while (m.kk >= 1 && m.cur.l == c) close_target(m, c0, c1, c2, bot_in, sb, cb, dst, s2b, act);
    }
  }
  if (m.kk >= 1) {
    if (!bot_in)
      m.acc += dp1c * a1;
    else
      m.acc = m.cur.d.z * pmean(c0, c1, c2, 0.f, sb);
  }
}

template <int B>
__global__ void __launch_bounds__(RF_COLS * RF_WAVES) k_remap_fields(const FieldArgs a) {
  extern __shared__ float cp_lds[];  // [nb][256] eliminated values at block ends
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // scalar: descriptors stay in SGPRs
  // blocks are dealt round-robin to the 8 XCDs: renumber so that the field
  // groups of one column block run on one XCD, one after another
  const unsigned bid = blockIdx.x, nx = 8, q8 = a.nwg / nx, r8 = a.nwg % nx;
  const unsigned xcd = bid % nx, idx = bid / nx;
  const unsigned logical = xcd < r8 ? xcd * (q8 + 1) + idx : r8 * (q8 + 1) + (xcd - r8) * q8 + idx;
  const int cblk = (int)(logical / a.nfg), fg = (int)(logical % a.nfg);
  const int f = fg * RF_WAVES + wave;
  if (f >= a.nf) return;  // whole wave; no barriers below
  const long long nc = a.ncol;
  const long long col_raw = (long long)cblk * RF_COLS + lane;
  const bool act = col_raw < nc;
  const long long col = act ? col_raw : nc - 1;
  const int ii = (int)(col % a.ni), jj = (int)(col / a.ni);
  const long long cell = (long long)(jj + a.js - a.jsd) * a.nid + (ii + a.is - a.isd);
  const long long s2 = a.s2;
  const int km = a.km;
  const bool is_pt = f >= a.nq;
  const int iv = is_pt ? 1 : 0;
  const int akord = is_pt ? a.kord_tm : a.kord_tr;
  const float qmin = is_pt ? a.t_min : 0.f;
  const unsigned s2b = (unsigned)s2 * 4u, fbytes = (unsigned)km * s2b;
  const Rsrc rsrc = make_rsrc(is_pt ? a.pt_src : a.q_src + (long long)f * km * s2, fbytes);
  const Rsrc rdst = make_rsrc(is_pt ? a.pt_dst : a.q_dst + (long long)f * km * s2, fbytes);
  const unsigned vq = (unsigned)cell * 4u;  // the lane's cell in a field level
  auto src = [&](int k) { return ldf(rsrc, vq, (unsigned)(k - 1) * s2b); };
  ColBuf cb;
  cb.r = make_rsrc(a.cols.d4, a.b_cols);
  cb.rc = make_rsrc(a.cols.cut, a.b_cut);
  cb.vc = (unsigned)col * 4u;
  cb.lvl = (unsigned)nc * 4u;
  cb.o_rbt = a.o_rbt; cb.o_gg = a.o_gg; cb.o_dp1 = a.o_dp1; cb.o_bot = a.o_bot; cb.o_lk = a.o_lk;
  float *cp = cp_lds + tid;

  // ---- downward elimination; keep the value at every block end
  float a_prev = src(1);
  float qp = (cb.d4(1) * a_prev + src(2)) * cb.rbt(1);
  for (int k0 = 2; k0 <= km; k0 += B) {
    float av[B], dv[B], rv[B];
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int k = min(k0 + j, km);
      av[j] = src(k);
      dv[j] = cb.d4(k);
      rv[j] = cb.rbt(k);
    }
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int k = k0 + j;
      if (k <= km) {
        qp = (3.f * (a_prev + dv[j] * av[j]) - qp) * rv[j];
        a_prev = av[j];
        if (k % B == 0 && k < km) cp[(k / B) * (RF_COLS * RF_WAVES)] = qp;
      }
    }
  }
  // qp = Q'(km); bottom row gives the raw interface value Q(km+1)
  const float a_km = a_prev, a_km1 = src(km - 1);
  const float qbot = (cb.bot(0) * a_km + a_km1 - cb.bot(1) * qp) * cb.bot(2);

  // ---- upward pass: window state after position p+1
  float A_m1 = a_km1, A_0 = a_km, A_p1 = 0.f, A_p2 = 0.f, A_p3 = 0.f;
  float qraw_next = qbot, QL1 = qbot, QL2 = 0.f;
  int E1 = 0, E2 = 0;
  const bool lim = akord <= 16;
  MapState ms;
  ms.kk = km;
  ms.acc = 0.f;
  ms.tvo = (unsigned)(km - 1) * cb.lvl + cb.vc;
  ms.dvo = (unsigned)(km - 1) * s2b + vq;
  ms.cur = load_target(cb, ms.tvo);

  for (int b = a.nb - 1; b >= 0; b--) {
    const int k0 = b * B + 1;
    const int cnt = min(B, km - k0 + 1);
    float av[B + 2];  // a1(k0-2 .. k0+B-1)
    float dv[B], rv[B], gv[B], pv[B];
#pragma unroll
    for (int j = 0; j < B + 2; j++) {
      const int k = min(max(k0 - 2 + j, 1), km);
      av[j] = src(k);
    }
#pragma unroll
    for (int j = 0; j < B; j++) {
      const int k = min(k0 + j, km);
      dv[j] = cb.d4(k);
      rv[j] = cb.rbt(k);
      gv[j] = cb.gg(k);
      pv[j] = cb.dp1(k);
    }
    // regenerate Q'(k0 .. k0+cnt-1)
    float qv[B];
    if (b == 0) {
      qv[0] = (dv[0] * av[2] + av[3]) * rv[0];
    } else {
      const float q_in = cp[b * (RF_COLS * RF_WAVES)];
      qv[0] = (3.f * (av[1] + dv[0] * av[2]) - q_in) * rv[0];
    }
#pragma unroll
    for (int j = 1; j < B; j++) qv[j] = (3.f * (av[j + 1] + dv[j] * av[j + 2]) - qv[j - 1]) * rv[j];

#pragma unroll
    for (int j = B - 1; j >= 0; j--) {
      if (j < cnt) {
        const int p = k0 + j;
        const float A_m2 = av[j];  // a1(p-2) (clamped, unused when p <= 2)
        const float qr = qv[j] - gv[j] * qraw_next;
        float ql = qr;
        if (lim) {
          const float lo = fminf(A_m1, A_0), hi = fmaxf(A_m1, A_0);
          if (p == 2 || p == km) {
            ql = fmaxf(fminf(ql, hi), lo);
          } else if (p >= 3) {
            const float glo = A_m1 - A_m2, ghi = A_p1 - A_0;
            if (glo * ghi > 0.f) {
              ql = fmaxf(fminf(ql, hi), lo);
            } else if (glo > 0.f) {
              ql = fmaxf(ql, lo);
            } else {
              ql = fminf(ql, hi);
              if (iv == 0) ql = fmaxf(0.f, ql);
            }
          }
        }
        // flags of cell p
        int e0;
        {
          const bool em = (p == 1 || p == km) ? (ql - A_0) * (QL1 - A_0) > 0.f
                                              : (A_0 - A_m1) * (A_p1 - A_0) < 0.f;
          e0 = em ? 1 : 0;
          if (akord > 9) {
            const float x0 = 2.f * A_0 - (ql + QL1), x1 = fabsf(ql - QL1);
            e0 |= (fabsf(x0) > x1 ? 2 : 0) | (fabsf(3.f * x0) > x1 ? 4 : 0);
          }
        }
        // cell c = p + 1 is now final
        if (p + 1 <= km) {
          const int c = p + 1;
          float a1c = A_p1, a2 = QL1, a3 = QL2, a4;
          cell_limit(c, km, iv, akord, qmin, a1c, a2, a3, a4, e0, E1, E2, A_0 - A_m1, A_p1 - A_0,
                     A_p2 - A_p1, A_p3 - A_p2);
          const float dpc = j + 1 < B ? pv[j + 1] : cb.dp1(c);
          map_layer(ms, c, km, A_p1, a2, a3, a4, dpc, cb, rdst, s2b, act);
        }
        qraw_next = qr;
        QL2 = QL1;
        QL1 = ql;
        E2 = E1;
        E1 = e0;
        A_p3 = A_p2;
        A_p2 = A_p1;
        A_p1 = A_0;
        A_0 = A_m1;
        A_m1 = A_m2;
      }
    }
  }
  // top cell: a2 = Ql(1), a3 = Ql(2); window now holds a1(1) in A_p1
  {
    float a1 = A_p1, a2 = QL1, a3 = QL2, a4;
    if (lim) {
      if (iv == 0) a2 = fmaxf(0.f, a2);
      a4 = 3.f * (2.f * a1 - (a2 + a3));
      cs_lim1<float>(a1, a2, a3, a4, (E1 & 1) != 0, 1);
    } else {
      a4 = 3.f * (2.f * a1 - (a2 + a3));
    }
    map_layer(ms, 1, km, a1, a2, a3, a4, cb.dp1(1), cb, rdst, s2b, act);
  }
}


// ---- levels across lanes ----------------------------------------------------
// k_remap_rows<M, L>: a wave holds 64/L columns, L lanes each (one DPP row for
// L = 16, two for L = 32); lane g of a column holds its levels g*M+1 ..
// g*M+M (those cells, and the interface at the top of each).  The column
// data (elimination factors and the target cut list) stay in registers while
// the wave walks every field of its columns, so each field is read from HBM
// once:
//   - the interface solve is the same Thomas recurrence, run inside each lane
//     from a zero entry value, composed across the column by an affine scan
//     (the slopes are field independent and scanned once per column), and run
//     again inside each lane from the true entry value;
//   - the interface limits, extremum flags and cell limiters are evaluated per
//     level with the neighbouring lanes' values moved in by DPP; the first two
//     and last two cells and interfaces take their own forms;
//   - each cell's parabola and mass go to LDS, and each lane integrates its
//     Eulerian layers from the parabolas of the source layers that hold their
//     top and bottom interfaces, plus the masses of the layers in between.
// The pieces of each Eulerian layer are summed in the order of the
// one-thread-per-(column, field) kernel above (bottom piece, whole layers
// upward, top piece).
constexpr int RW_WAVES = 8;  // waves per workgroup

struct RowsArgs {
  int km, ni, is, js, isd, jsd, nid, nq, nf, kord_tr, kord_tm;
  unsigned nwg;
  long long ncol, s2;
  float t_min;
  const float *q_src, *pt_src;
  float *q_dst, *pt_dst;
  RemapCols cols;
};

// Cross-lane moves.  Each result is pinned by an empty volatile asm where it
// is computed: a DPP or bpermute that the compiler moved into a lane-dependent
// branch would run with some source lanes disabled and read 0 from them.
__device__ __forceinline__ int pin(int v) {
  asm volatile("" : "+v"(v));
  return v;
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp0(int v) {  // 0 where the source lane is invalid
  return pin(__builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ float dpp0f(float v) {
  return __builtin_bit_cast(float, dpp0<CTRL>(__builtin_bit_cast(int, v)));
}
// lane l reads lane l-N (l+N) of its 16-lane row; 0 past the row's end
template <int N>
__device__ __forceinline__ float row_lo(float v) { return dpp0f<0x110 + N>(v); }
template <int N>
__device__ __forceinline__ float row_hi(float v) { return dpp0f<0x100 + N>(v); }
// the previous / next lane of the column: within the row (L = 16), or across
// the whole wave (L = 32; at a column boundary it reads the other column's
// lane, whose value is never used there); 0 past the wave's end
template <int L>
__device__ __forceinline__ int nb_lo_i(int v) { return dpp0<L == 16 ? 0x111 : 0x138>(v); }
template <int L>
__device__ __forceinline__ int nb_hi_i(int v) { return dpp0<L == 16 ? 0x101 : 0x130>(v); }
template <int L>
__device__ __forceinline__ float nb_lo(float v) { return __builtin_bit_cast(float, nb_lo_i<L>(__builtin_bit_cast(int, v))); }
template <int L>
__device__ __forceinline__ float nb_hi(float v) { return __builtin_bit_cast(float, nb_hi_i<L>(__builtin_bit_cast(int, v))); }

// inclusive affine scans over the lanes of a column: lane g holds the map
// y -> s y + v of its levels; after the scan v is the map's value for a zero
// entry at the column's end (lo: from lane 0 upward; hi: from lane L-1
// downward).  s[j] are the lanes' slope products at each step.
template <int L>
__device__ __forceinline__ float scan_lo(float v, const float *s) {
  v = fmaf(s[0], row_lo<1>(v), v);
  v = fmaf(s[1], row_lo<2>(v), v);
  v = fmaf(s[2], row_lo<4>(v), v);
  v = fmaf(s[3], row_lo<8>(v), v);
  if (L == 32)  // the second row of each column takes the first row's last lane (row_bcast:15)
    v = fmaf(s[4], __builtin_bit_cast(float, dpp0<0x142, 0xa>(__builtin_bit_cast(int, v))), v);
  return v;
}
template <int L>
__device__ __forceinline__ float scan_hi(float v, const float *s, int lane) {
  v = fmaf(s[0], row_hi<1>(v), v);
  v = fmaf(s[1], row_hi<2>(v), v);
  v = fmaf(s[2], row_hi<4>(v), v);
  v = fmaf(s[3], row_hi<8>(v), v);
  if (L == 32) {  // the first row of each column takes the second row's first lane
    const float u = __builtin_bit_cast(float, pin(__builtin_amdgcn_ds_bpermute(((lane & 32) | 16) << 2, __builtin_bit_cast(int, v))));
    v = (lane & 16) == 0 ? fmaf(s[4], u, v) : v;
  }
  return v;
}
// slope products: a lane with no partner at a step keeps its product (the
// second row of a column needs its own row's product at the cross-row step)
template <int CTRL>
__device__ __forceinline__ float slope_mv(float v) {
  return __builtin_bit_cast(float, pin(__builtin_amdgcn_update_dpp(__builtin_bit_cast(int, 1.f), __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false)));
}
__device__ __forceinline__ void slopes_lo(float sl, float *s) {
  s[0] = sl; sl *= slope_mv<0x111>(sl);
  s[1] = sl; sl *= slope_mv<0x112>(sl);
  s[2] = sl; sl *= slope_mv<0x114>(sl);
  s[3] = sl; sl *= slope_mv<0x118>(sl);
  s[4] = sl;
}
__device__ __forceinline__ void slopes_hi(float sl, float *s) {
  s[0] = sl; sl *= slope_mv<0x101>(sl);
  s[1] = sl; sl *= slope_mv<0x102>(sl);
  s[2] = sl; sl *= slope_mv<0x104>(sl);
  s[3] = sl; sl *= slope_mv<0x108>(sl);
  s[4] = sl;
}

// staging tile of a field: [column of the workgroup][level], the level index
// skewed by k >> 5 and the row padded to an odd length, so that the
// coalesced writes (consecutive columns) and the per-lane reads (M levels
// 4 apart across lanes) spread over the LDS banks
template <int LM>
struct Stage {
  static constexpr int row = LM + LM / 32 + 1 + ((LM + LM / 32 + 1) % 2 == 0 ? 1 : 0);
  static __device__ __forceinline__ int at(int c, int k) { return c * row + k + (k >> 5); }
};

// cell records: [column][level], the level index k stored at k + (k >> 2),
// so that the lanes' 16-byte writes (4 levels apart) fall on distinct bank
// groups for 8 or 16 lanes at a time
template <int LM>
struct Recs {
  static constexpr int row = LM + LM / 4;
  static __device__ __forceinline__ int at(int c, int k) { return c * row + k + (k >> 2); }
};

template <int M, int L>
size_t rows_lds_bytes() {
  constexpr int wgc = RW_WAVES * (64 / L);
  return sizeof(float4) * wgc * Recs<L * M>::row + 4 * sizeof(float) * wgc * Stage<L * M>::row;
}

template <int M, int L>
__global__ void __launch_bounds__(64 * RW_WAVES, 4) k_remap_rows(const RowsArgs a) {
  constexpr int CPW = 64 / L;              // columns per wave
  constexpr int WGC = RW_WAVES * CPW;      // columns per workgroup
  constexpr int NT = 64 * RW_WAVES;        // threads per workgroup
  constexpr int KS = NT / WGC;             // level step of a thread's staged elements
  using ST = Stage<L * M>;
  using RC = Recs<L * M>;
  // LDS: cell records (c0 c1 c2 mass) of the workgroup's columns, then the
  // input and output staging tiles of two consecutive fields (one barrier per
  // field: field f's output tile is stored after field f+1's barrier)
  extern __shared__ float4 rw_lds[];
  float *const s_tiles = reinterpret_cast<float *>(rw_lds + WGC * RC::row);
  auto tile_in = [&](int f) { return s_tiles + (f & 1) * 2 * WGC * ST::row; };
  auto tile_out = [&](int f) { return s_tiles + ((f & 1) * 2 + 1) * WGC * ST::row; };
  auto field_dst = [&](int f) { return f >= a.nq ? a.pt_dst : a.q_dst + (long long)f * a.km * a.s2; };
  const int tid = threadIdx.x, lane = tid & 63, cw = lane / L;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  // blocks are dealt round-robin to the 8 XCDs: renumber so that neighbouring
  // column blocks (which share cache lines) run on one XCD
  const unsigned bid = blockIdx.x, nx = 8, q8 = a.nwg / nx, r8 = a.nwg % nx;
  const unsigned xcd = bid % nx, idx = bid / nx;
  const unsigned logical = xcd < r8 ? xcd * (q8 + 1) + idx : r8 * (q8 + 1) + (xcd - r8) * q8 + idx;
  const long long nc = a.ncol;
  const long long wcol0 = (long long)logical * WGC;  // the workgroup's first column
  // (columns past the last one compute on a copy of it and store nothing;
  // every wave takes part in the workgroup barriers)
  const int wc = wave * CPW + cw;  // the lane's column within the workgroup
  const long long col = std::min(wcol0 + wc, nc - 1);
  const int km = a.km;
  const int gc = lane & (L - 1);
  const int k0c = gc * M;  // the lane's levels are k0+1 .. k0+M
  const unsigned s2b = (unsigned)a.s2 * 4u, fbytes = (unsigned)km * s2b;
  // the thread's staged elements: column tc of the workgroup, levels
  // tk0 + j KS (j < M), read and written with consecutive columns across
  // lanes; levels past km fall past the buffer end, where the hardware reads 0
  // and drops stores; stores of columns past the last go there too
  const int tc = tid % WGC, tk0 = tid / WGC;
  unsigned vtc, vts;
  {
    const bool t_act = wcol0 + tc < nc;
    const long long tcol = t_act ? wcol0 + tc : nc - 1;
    const int ti = (int)(tcol % a.ni), tj = (int)(tcol / a.ni);
    const unsigned tcell = (unsigned)((tj + a.js - a.jsd) * a.nid + (ti + a.is - a.isd));
    vtc = (unsigned)tk0 * s2b + tcell * 4u;
    vts = t_act ? vtc : 0x80000000u;
  }
  const unsigned kstep = (unsigned)KS * s2b;
  // LDS bases: the thread's staged elements sit at compile-time offsets
  // from sbt (its levels are tk0 + j KS with tk0 < KS <= 32), the lane's
  // levels k0 + i at offsets i from slr and srec (k0 is a multiple of M and
  // M <= 4, so neither skew changes across them)
  static_assert(KS <= 32 && (KS & (KS - 1)) == 0 && M <= 4, "staging offsets");
  const int sbt = ST::at(tc, tk0), slr = ST::at(wc, k0c), srec = RC::at(wc, k0c);

  // ---- column data (levels past km read 0)
  const unsigned lvl = (unsigned)nc * 4u, cbytes = (unsigned)km * lvl;
  float d4[M], rbt[M], gg[M], dp1[M], st[M], wt[M], wb[M], rd[M];
  int ltc[M];
  {
    const Rsrc r_d4 = make_rsrc(a.cols.d4, cbytes), r_rbt = make_rsrc(a.cols.rbt, cbytes);
    const Rsrc r_gg = make_rsrc(a.cols.gg, cbytes), r_dp1 = make_rsrc(a.cols.dp1, cbytes);
    const Rsrc r_lk = make_rsrc(a.cols.lk, cbytes), r_cut = make_rsrc(a.cols.cut, 4u * cbytes);
    const unsigned vc = (unsigned)k0c * lvl + (unsigned)col * 4u;
#pragma unroll
    for (int i = 0; i < M; i++) {
      const unsigned o = vc + (unsigned)i * lvl;
      d4[i] = ldf(r_d4, o, 0);
      rbt[i] = ldf(r_rbt, o, 0);
      gg[i] = ldf(r_gg, o, 0);
      dp1[i] = ldf(r_dp1, o, 0);
      ltc[i] = ldi(r_lk, o, 0);
      st[i] = ldf(r_cut, 4u * o, 0);
      wt[i] = ldf(r_cut, 4u * o + 4u, 0);
      wb[i] = ldf(r_cut, 4u * o + 8u, 0);
      rd[i] = ldf(r_cut, 4u * o + 12u, 0);
      // interface km+1 (and the padding past it) lies at the bottom of layer km
      if (k0c + i + 1 > km) {
        ltc[i] = km;
        st[i] = 1.f;
      }
    }
  }
  const float bot0 = a.cols.bot[col], bot1 = a.cols.bot[nc + col], bot2 = a.cols.bot[2 * nc + col];
  // the source layer and fraction of the bottom interface of the lane's last
  // target (the last lane's last target is padding: km+1 <= L M)
  // (the moves run for every lane before the selection: the last lane's value
  // is read by its neighbour)
  const int ltn_n = nb_hi_i<L>(ltc[0]);
  const float stn_n = nb_hi<L>(st[0]);
  const int ltn = gc == L - 1 ? km : ltn_n;
  const float stn = gc == L - 1 ? 1.f : stn_n;
  // LDS indices of the records of those source layers
  int ot[M];
#pragma unroll
  for (int i = 0; i < M; i++) ot[i] = RC::at(wc, ltc[i] - 1);
  const int otn = RC::at(wc, ltn - 1);
  // slopes of the lane's forward (Q'(k0) -> Q'(k0+M)) and backward
  // (Q(k0+M+1) -> Q(k0+1)) maps, and their products at the scan steps
  float fs[5], bs[5];
  {
    float sf = 1.f, sb = 1.f;
#pragma unroll
    for (int i = 0; i < M; i++) {
      sf *= -rbt[i];
      sb *= -gg[i];
    }
    slopes_lo(sf, fs);
    slopes_hi(sb, bs);
  }
  // wave-uniform positions of the special levels
  const int g_km = (km - 1) / M, s_km = (km - 1) % M;      // cell / interface km
  const int g_km1 = (km - 2) / M, s_km1 = (km - 2) % M;    // cell km-1
  const int g_kb = km / M, s_kb = km % M;                  // interface km+1

  float qn[M];  // the thread's staged elements of the next field
  {
    const Rsrc r0 = make_rsrc(a.nq > 0 ? a.q_src : a.pt_src, fbytes);
#pragma unroll
    for (int j = 0; j < M; j++) qn[j] = ldf(r0, vtc + (unsigned)j * kstep, 0);
  }
  for (int f = 0; f < a.nf; f++) {
    // level-dependent offsets, masks and LDS addresses are cheaper to
    // recompute every field than to keep live across the loop
    int g = gc;
    asm volatile("" : "+v"(g));
    auto rec_of = [&](int l) -> float4 & { return rw_lds[RC::at(wc, l - 1)]; };
    const int k0 = k0c;

    const bool is_pt = f >= a.nq;
    const int iv = is_pt ? 1 : 0;
    const int akord = is_pt ? a.kord_tm : a.kord_tr;
    const float qmin = is_pt ? a.t_min : 0.f;
    const bool lim = akord <= 16;
    // stage the field; then each lane takes its column's levels
#pragma unroll
    for (int j = 0; j < M; j++) tile_in(f)[sbt + j * KS + ((j * KS) >> 5)] = qn[j];
    __syncthreads();
    if (f > 0) {  // the previous field's results, staged before the barrier
      const Rsrc rprev = make_rsrc(field_dst(f - 1), fbytes);
#pragma unroll
      for (int j = 0; j < M; j++) stf(rprev, vts + (unsigned)j * kstep, tile_out(f - 1)[sbt + j * KS + ((j * KS) >> 5)]);
    }
    if (f + 1 < a.nf) {  // next field's elements, in flight while this one is computed
      const int fn = f + 1;
      const float *src = fn >= a.nq ? a.pt_src : a.q_src + (long long)fn * km * a.s2;
#pragma unroll
      for (int j = 0; j < M; j++) qn[j] = ldf(make_rsrc(src, fbytes), vtc + (unsigned)j * kstep, 0);
    }
    float q[M];
#pragma unroll
    for (int i = 0; i < M; i++) q[i] = tile_in(f)[slr + i];

    // cell means of levels k0-1 .. k0+M+2 (A(j) is level k0+j+1)
    const float am2 = nb_lo<L>(q[M - 2]), am1 = nb_lo<L>(q[M - 1]);
    const float ap0 = nb_hi<L>(q[0]), ap1 = nb_hi<L>(q[1]);
    auto A = [&](int j) -> float { return j == -2 ? am2 : j == -1 ? am1 : j == M ? ap0 : j == M + 1 ? ap1 : q[j]; };

    // ---- forward elimination: Q'(k) = (3 (a(k-1) + d4 a(k)) - Q'(k-1)) rbt;
    // level 1: Q'(1) = (d4 a(1) + a(2)) rbt
    const float t1 = fmaf(d4[0], q[0], q[1]);
    float x[M];
    {
      float e = (g == 0 ? t1 : 3.f * fmaf(d4[0], q[0], am1)) * rbt[0];
#pragma unroll
      for (int i = 1; i < M; i++) e = (3.f * fmaf(d4[i], q[i], q[i - 1]) - e) * rbt[i];
      const float cin = nb_lo<L>(scan_lo<L>(e, fs));
      e = g == 0 ? t1 * rbt[0] : (3.f * fmaf(d4[0], q[0], am1) - cin) * rbt[0];
      x[0] = e;
#pragma unroll
      for (int i = 1; i < M; i++) x[i] = e = (3.f * fmaf(d4[i], q[i], q[i - 1]) - e) * rbt[i];
    }
    // bottom row: Q(km+1) from a(km), a(km-1) and Q'(km), placed at interface km+1
    {
      float a_km = 0.f, a_km1 = 0.f, x_km = 0.f;
#pragma unroll
      for (int i = 0; i < M; i++)
        if (i == s_km) {
          a_km = q[i];
          a_km1 = A(i - 1);
          x_km = x[i];
        }
      float qb = (bot0 * a_km + a_km1 - bot1 * x_km) * bot2;
      if (s_kb == 0) qb = nb_lo<L>(qb);
#pragma unroll
      for (int i = 0; i < M; i++)
        if (i == s_kb && g == g_kb) x[i] = qb;
    }
    // ---- back substitution: Q(k) = Q'(k) - gg Q(k+1) (gg is 0 from km+1 on)
    {
      float e = 0.f;
#pragma unroll
      for (int i = M - 1; i >= 0; i--) e = fmaf(-gg[i], e, x[i]);
      e = nb_hi<L>(scan_hi<L>(e, bs, lane));
#pragma unroll
      for (int i = M - 1; i >= 0; i--) x[i] = e = fmaf(-gg[i], e, x[i]);
    }

    // ---- interface limits, in place (interface p = k0+i+1 lies between cells
    // p-1 and p); interfaces 1 and km+1 take no limit, 2 and km the two-sided
    // one (their slots are known to the whole wave: only there do the lanes
    // that hold them take their own forms)
    if (lim) {
      const float dn_floor = iv == 0 ? 0.f : -INFINITY;
#pragma unroll
      for (int i = 0; i < M; i++) {
        const float am = A(i - 1), a0 = A(i);
        const float lo = fminf(am, a0), hi = fmaxf(am, a0);
        const float glo = am - A(i - 2), ghi = A(i + 1) - a0;
        const float both = __builtin_amdgcn_fmed3f(x[i], lo, hi);  // lo <= hi: the clamp
        const float up = fmaxf(x[i], lo);
        const float dn = fmaxf(fminf(x[i], hi), dn_floor);
        float v = glo * ghi > 0.f ? both : (glo > 0.f ? up : dn);
        if (i == 0) {
          if (g == 0) v = x[i];
        }
        if (i == 1) {
          if (g == 0) v = both;
        }
        if (i == s_km) {
          if (g == g_km) v = both;
        }
        if (i == s_kb) {
          if (g == g_kb) v = x[i];
        }
        x[i] = v;
      }
    }
    const float qln = nb_hi<L>(x[0]);  // interface k0+M+1
    auto QL = [&](int j) -> float { return j == M ? qln : x[j]; };

    // ---- profile of every cell (a2 = Ql(k), a3 = Ql(k+1), a4), staged in LDS
    // with the cell's mass
    if (lim) {
      unsigned eb = 0;  // 3 flag bits per cell k0+i+1: 1 extm, 2 ext5, 4 ext6
#pragma unroll
      for (int i = 0; i < M; i++) {
        const float a0 = A(i);
        bool em = (a0 - A(i - 1)) * (A(i + 1) - a0) < 0.f;
        // cells 1 and km: extremum against their own interface values
        if (i == 0) {
          if (g == 0) em = (QL(i) - a0) * (QL(i + 1) - a0) > 0.f;
        }
        if (i == s_km) {
          if (g == g_km) em = (QL(i) - a0) * (QL(i + 1) - a0) > 0.f;
        }
        unsigned e = em ? 1u : 0u;
        if (akord > 9) {
          const float x0 = 2.f * a0 - (QL(i) + QL(i + 1)), x1 = fabsf(QL(i) - QL(i + 1));
          e |= (fabsf(x0) > x1 ? 2u : 0u) | (fabsf(3.f * x0) > x1 ? 4u : 0u);
        }
        eb |= e << (3 * i);
      }
      const int e_prev = (int)((unsigned)nb_lo_i<L>((int)eb) >> (3 * (M - 1))) & 7;
      const int e_next = nb_hi_i<L>((int)eb) & 7;
      auto EF = [&](int j) -> int { return j < 0 ? e_prev : j >= M ? e_next : (int)(eb >> (3 * j)) & 7; };
#pragma unroll
      for (int i = 0; i < M; i++) {
        float a1 = q[i], a2 = QL(i), a3 = QL(i + 1), a4;
        cell_limit_interior(iv, akord, qmin, a1, a2, a3, a4, EF(i - 1), EF(i), EF(i + 1), A(i - 1) - A(i - 2),
                            A(i) - A(i - 1), A(i + 1) - A(i), A(i + 2) - A(i + 1));
        // cells 1, 2, km-1 and km take their own forms (at slots the whole
        // wave knows, in the lanes that hold them)
        // (lmt: cs_limiters mode; clip: for tracers the top (1) or bottom (2)
        // interface value is bounded below by 0)
        auto edge = [&](int lmt, int clip) {
          float b1 = a1, b2 = QL(i), b3 = QL(i + 1);
          if (iv == 0 && clip == 1) b2 = fmaxf(0.f, b2);
          if (iv == 0 && clip == 2) b3 = fmaxf(0.f, b3);
          float b4 = 3.f * (2.f * b1 - (b2 + b3));
          cs_lim1<float>(b1, b2, b3, b4, (EF(i) & 1) != 0, lmt);
          a2 = b2;
          a3 = b3;
          a4 = b4;
        };
        if (i == 0) {
          if (g == 0) edge(1, 1);  // top cell
        }
        if (i == 1) {
          if (g == 0) edge(2, 0);
        }
        if (i == s_km1) {
          if (g == g_km1) edge(2, 0);
        }
        if (i == s_km) {
          if (g == g_km) edge(1, 2);  // bottom cell
        }
        rw_lds[srec + i] = make_float4(a2, 0.5f * (a4 + a3 - a2), a4 * (1.0f / 3.0f), dp1[i] * q[i]);
      }
    } else {
#pragma unroll
      for (int i = 0; i < M; i++) {
        const float a2 = QL(i), a3 = QL(i + 1), a4 = 3.f * (2.f * q[i] - (a2 + a3));
        rw_lds[srec + i] = make_float4(a2, 0.5f * (a4 + a3 - a2), a4 * (1.0f / 3.0f), dp1[i] * q[i]);
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    // ---- Eulerian layers k0+1 .. k0+M
#pragma unroll
    for (int i = 0; i < M; i++) {
      const int l_t = ltc[i], l_b = i + 1 < M ? ltc[i + 1] : ltn;
      const float s_t = st[i], s_b = i + 1 < M ? st[i + 1] : stn;
      const float4 ct = rw_lds[ot[i]], cb = rw_lds[i + 1 < M ? ot[i + 1] : otn];
      // within one source layer: the parabola's mean over [s_t, s_b]; else
      // the bottom piece, the whole layers between, and the top piece over
      // [s_t, 1], weighted by their pressure widths
      const bool same = l_t == l_b;
      const float top = pmean(ct.x, ct.y, ct.z, s_t, same ? s_b : 1.f);
      float acc = wb[i] * pmean(cb.x, cb.y, cb.z, 0.f, s_b);
      for (int l = l_b - 1; l > l_t; l--) acc += rec_of(l).w;
      tile_out(f)[slr + i] = same ? top : (acc + wt[i] * top) * rd[i];
    }
  }
  __syncthreads();
  {
    const int f = a.nf - 1;
    const Rsrc rlast = make_rsrc(field_dst(f), fbytes);
#pragma unroll
    for (int j = 0; j < M; j++) stf(rlast, vts + (unsigned)j * kstep, tile_out(f)[sbt + j * KS + ((j * KS) >> 5)]);
  }
}

}  // namespace

size_t remap_cols_bytes(int km, long long ncol) {
  const size_t per = (size_t)ncol;
  size_t b = 0;
  auto add = [&](size_t n) { b += (n + 255) & ~(size_t)255; };
  for (int i = 0; i < 4; i++) add(sizeof(float) * km * per);  // d4 rbt gg dp1
  add(sizeof(float) * 3 * per);                                // bot
  add(sizeof(int) * km * per);                                 // lk
  add(sizeof(float4) * km * per);                              // cut
  return b;
}

RemapCols remap_cols_carve(void *base, int km, long long ncol) {
  RemapCols c{};
  char *p = (char *)base;
  const size_t per = (size_t)ncol;
  auto take = [&](size_t n) {
    char *r = p;
    p += (n + 255) & ~(size_t)255;
    return r;
  };
  c.d4 = (float *)take(sizeof(float) * km * per);
  c.rbt = (float *)take(sizeof(float) * km * per);
  c.gg = (float *)take(sizeof(float) * km * per);
  c.dp1 = (float *)take(sizeof(float) * km * per);
  c.bot = (float *)take(sizeof(float) * 3 * per);
  c.lk = (int *)take(sizeof(int) * km * per);
  c.cut = (float4 *)take(sizeof(float4) * km * per);
  return c;
}

// the field kernel addresses a tile's column data (d4 .. lk, and cut) and one
// field's levels with 32-bit byte offsets
static int check_sizes(int km, long long ncol, long long s2) {
  auto al = [](long long b) { return (b + 255) & ~255LL; };
  const long long span = 4 * al(4LL * km * ncol) + al(12LL * ncol) + 4LL * km * ncol;
  if (span >= (1LL << 31) || 16LL * km * ncol >= (1LL << 31) || 4LL * km * s2 >= (1LL << 31))
    return set_error("fast remap: a tile's column data or field level set exceeds 2 GiB "
                     "(the one-thread-per-column remap, fv3_ctx_set_launch remap = 1, has no such limit)");
  return 0;
}

// the levels-across-lanes kernel addresses up to 128 levels (the padding past
// km included) of a field and of the column data with 32-bit byte offsets
static bool rows_fits(int km, long long ncol, long long s2) {
  (void)km;
  return 4LL * 128 * s2 < (1LL << 31) && 16LL * 128 * ncol < (1LL << 31);
}

static int check_orders(int km, int kord_tr, int kord_tm, bool with_pt) {
  if (km < 8) return set_error("fast remap needs km >= 8 (got %d)", km);
  if (!(kord_tr > 7 && (kord_tr <= 10 || kord_tr > 16)))
    return set_error("kord_tr=%d: implemented orders are 8, 9, 10 and > 16", kord_tr);
  if (with_pt && !(kord_tm > 7 && (kord_tm <= 10 || kord_tm > 16)))
    return set_error("kord_tm=%d: implemented orders are 8, 9, 10 and > 16", kord_tm);
  return 0;
}

int remap_fast_prep_f32(Ctx &c, int km, const double *ak, const double *bk, double ptop,
                        const double *dp1, double *delp_out, const RemapCols &cols, hipStream_t s) {
  const Dims &d = c.d;
  if (km < 8) return set_error("fast remap needs km >= 8 (got %d)", km);
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  const long long ncol = (long long)ni * nj;
  PrepArgs pa{};
  pa.km = km; pa.ni = ni; pa.is = d.is; pa.js = d.js; pa.isd = d.isd; pa.jsd = d.jsd; pa.nid = d.nid;
  pa.ncol = ncol; pa.s2 = d.s2; pa.ptop = ptop; pa.dp1 = dp1; pa.ak = ak; pa.bk = bk;
  pa.delp_out = delp_out; pa.cols = cols;
  hipLaunchKernelGGL(k_remap_prep, dim3((unsigned)((ncol + 255) / 256)), dim3(256), 0, s, pa);
  FV3_HIP(hipGetLastError());
  return 0;
}

// More than 64 KB of dynamic LDS per workgroup must be allowed per kernel:
// done once per process, at context creation (before any stream capture).
int remap_rows_setup() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
    const void *fn[3] = {reinterpret_cast<const void *>(&k_remap_rows<2, 16>),
                         reinterpret_cast<const void *>(&k_remap_rows<4, 16>),
                         reinterpret_cast<const void *>(&k_remap_rows<4, 32>)};
    const size_t lds[3] = {rows_lds_bytes<2, 16>(), rows_lds_bytes<4, 16>(), rows_lds_bytes<4, 32>()};
    for (int v = 0; v < 3 && err == hipSuccess; v++)
      err = hipFuncSetAttribute(fn[v], hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds[v]);
  });
  FV3_HIP(err);
  return 0;
}

int remap_fast_fields_f32(Ctx &c, int km, int nq, const float *q_src, float *q_dst, const float *pt_src,
                          float *pt_dst, int kord_tr, int kord_tm, double t_min, const RemapCols &cols,
                          hipStream_t s) {
  const Dims &d = c.d;
  if (check_orders(km, kord_tr, kord_tm, pt_src != nullptr)) return 1;
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  const long long ncol = (long long)ni * nj;
  FieldArgs fa{};
  fa.km = km; fa.nb = (km + RF_B - 1) / RF_B; fa.ni = ni; fa.is = d.is; fa.js = d.js;
  fa.isd = d.isd; fa.jsd = d.jsd; fa.nid = d.nid; fa.nq = nq; fa.nf = nq + (pt_src ? 1 : 0);
  fa.nfg = (fa.nf + RF_WAVES - 1) / RF_WAVES; fa.kord_tr = kord_tr; fa.kord_tm = kord_tm;
  fa.ncol = ncol; fa.s2 = d.s2; fa.t_min = (float)t_min;
  fa.q_src = q_src; fa.pt_src = pt_src; fa.q_dst = q_dst; fa.pt_dst = pt_dst; fa.cols = cols;
  {
    if (check_sizes(km, ncol, d.s2)) return 1;
    const char *b = (const char *)cols.d4;
    const long long cut = (long long)sizeof(float4) * km * ncol, span = (const char *)cols.lk - b + 4LL * km * ncol;
    fa.o_rbt = (unsigned)((const char *)cols.rbt - b);
    fa.o_gg = (unsigned)((const char *)cols.gg - b);
    fa.o_dp1 = (unsigned)((const char *)cols.dp1 - b);
    fa.o_bot = (unsigned)((const char *)cols.bot - b);
    fa.o_lk = (unsigned)((const char *)cols.lk - b);
    fa.b_cols = (unsigned)span;
    fa.b_cut = (unsigned)cut;
  }
  if (c.remap_mode == 0 && km + 1 <= 128 && rows_fits(km, ncol, d.s2)) {
    RowsArgs ra{};
    ra.km = km; ra.ni = ni; ra.is = d.is; ra.js = d.js; ra.isd = d.isd; ra.jsd = d.jsd; ra.nid = d.nid;
    ra.nq = nq; ra.nf = fa.nf; ra.kord_tr = kord_tr; ra.kord_tm = kord_tm;
    ra.ncol = ncol; ra.s2 = d.s2; ra.t_min = (float)t_min;
    ra.q_src = q_src; ra.pt_src = pt_src; ra.q_dst = q_dst; ra.pt_dst = pt_dst; ra.cols = cols;
    // levels per lane x lanes per column: 2 x 16, 4 x 16 or 4 x 32
    const int M = km + 1 <= 32 ? 2 : 4, L = km + 1 <= 64 ? 16 : 32;
    const long long per_wg = (long long)RW_WAVES * (64 / L);
    const long long nwg = (ncol + per_wg - 1) / per_wg;
    if (nwg > 0x7fffffffLL) return set_error("remap grid too large");
    ra.nwg = (unsigned)nwg;
    const size_t lds = M == 2 ? rows_lds_bytes<2, 16>() : L == 16 ? rows_lds_bytes<4, 16>() : rows_lds_bytes<4, 32>();
    if (remap_rows_setup()) return 1;
    if (M == 2)
      hipLaunchKernelGGL((k_remap_rows<2, 16>), dim3((unsigned)nwg), dim3(64 * RW_WAVES), lds, s, ra);
    else if (L == 16)
      hipLaunchKernelGGL((k_remap_rows<4, 16>), dim3((unsigned)nwg), dim3(64 * RW_WAVES), lds, s, ra);
    else
      hipLaunchKernelGGL((k_remap_rows<4, 32>), dim3((unsigned)nwg), dim3(64 * RW_WAVES), lds, s, ra);
    FV3_HIP(hipGetLastError());
    return 0;
  }
  const long long ncb = (ncol + RF_COLS - 1) / RF_COLS;
  const long long nwg = ncb * fa.nfg;
  if (nwg > 0x7fffffffLL) return set_error("remap grid too large");
  fa.nwg = (unsigned)nwg;
  const size_t lds = sizeof(float) * fa.nb * RF_COLS * RF_WAVES;
  if (lds > 64 * 1024) return set_error("fast remap supports km <= %d (got %d)", 64 * RF_B - 1, km);
  hipLaunchKernelGGL((k_remap_fields<RF_B>), dim3((unsigned)nwg), dim3(RF_COLS * RF_WAVES), lds, s, fa);
  FV3_HIP(hipGetLastError());
  return 0;
}

int remap_fast_check(int km, long long ncol, long long s2, int kord_tr, int kord_tm, bool with_pt) {
  if (check_orders(km, kord_tr, kord_tm, with_pt)) return 1;
  return check_sizes(km, ncol, s2);
}

}  // namespace fv3
