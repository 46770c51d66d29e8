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

// Limited cell coefficients of source layer c (2 <= c <= km) from the window.
// e_lo/e_c/e_hi: flag bits (1 extm, 2 ext5, 4 ext6) of cells c-1, c, c+1.
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
