// fused.hip -- FP64 instantiation of the fused transport kernel and the
// device-resident advection + remap step driver.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "fused.h"
#include "fused_kernel.h"
#include "mapz_fast.h"

namespace fv3 {

template <int HO>
static void go64(const FusedArgs<double> &a, const KindStreams &s) {
  const dim3 b(FT_LANES * 4);
  int n[7];
  fused_grid_split(a, n);
  auto grid = [&](int c) { return dim3((unsigned)c, (unsigned)a.npz, (unsigned)a.ntiles); };
  if (n[0] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 0, 1>), grid(n[0]), b, 0, s.k[0], a);
  if (n[1] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 1, 1>), grid(n[1]), b, 0, s.k[1], a);
  if (n[2] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 2, 1>), grid(n[2]), b, 0, s.k[2], a);
  if (n[3] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 3, 1>), grid(n[3]), b, 0, s.k[2], a);
  if (n[4] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 4, 1>), grid(n[4]), b, 0, s.k[3], a);
  if (n[5] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 5, 1>), grid(n[5]), b, 0, s.k[2], a);
  if (n[6] > 0) hipLaunchKernelGGL((k_fused_tp<double, 4, 1, HO, 6, 1>), grid(n[6]), b, 0, s.k[3], a);
}

int launch_fused_f64(const FusedArgs<double> &a, const KindStreams &s) {
  if (a.hord == 8)
    go64<8>(a, s);
  else
    go64<0>(a, s);
  FV3_HIP(hipGetLastError());
  return 0;
}

__global__ void k_nsplit_check(const int *nsplit, int npz, int cap, int *flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int bad = 0;
    for (int k = 0; k < npz; k++) bad |= nsplit[k] > cap || nsplit[k] < 1;
    *flag = bad;
  }
}

template <typename T>
static int fused_step_t(Ctx &c, int nq, const T *q_in, T *q_out, T *pt, const double *dp1,
                        double *delp_out, const T *cx, const T *cy, const T *mfx, const T *mfy,
                        const double *ak, const double *bk, double ptop, int hord, int kord_tr,
                        int kord_tm, double t_min, int nsplit_cap, const StepArena &arena) {
  const Dims &d = c.d;
  const int npz = c.npz;
  hipStream_t s = c.stream;
  if (hord < 7 || hord > 13) return set_error("hord=%d: only the monotone orders 7..13 are implemented", hord);
  if (c.graph_safe && nsplit_cap <= 0)
    return set_error("graph-safe steps need nsplit_cap > 0 (no device-to-host read)");
  // the host-array entry points run the transport alone (c.transport_only)
  const bool remap_phase = !c.transport_only;
  const bool fast_remap = remap_phase && std::is_same<T, float>::value && c.remap_mode != 1 && npz >= 8;
  // the per-column remap synchronises the host with the stream, so only the
  // column-shared FP32 remap can be captured
  if (c.graph_safe && !fast_remap)
    return set_error("graph-safe steps need the FP32 form with the column-shared remap and npz >= 8");
  const bool ext = arena.base != nullptr;
  if (ext && !fast_remap)
    return set_error("the no-allocation step needs the FP32 form with the column-shared remap and npz >= 8");
  // head: cmax, frac, nsplit, flag (the context's staging arena, or the
  // front of the caller's arena)
  size_t off = 0;
  auto take = [&](size_t b) {
    size_t o = off;
    off += (b + 255) & ~(size_t)255;
    return o;
  };
  const size_t o_cmax = take(sizeof(double) * npz), o_frac = take(sizeof(T) * npz);
  const size_t o_ns = take(sizeof(int) * npz), o_flag = take(sizeof(int));
  const size_t head = off;
  const int nt = c.ntiles;
  const long long qsize = d.s2 * npz * nq * nt;  // all tiles
  const int ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  // the FP32 remap's scratch: q_tmp, the pt source copy, every tile's column data
  const size_t b_q = ((sizeof(T) * qsize) + 255) & ~(size_t)255;
  const size_t b_pt = pt ? ((sizeof(T) * d.s2 * npz * nt) + 255) & ~(size_t)255 : 0;
  const size_t b_col = fast_remap ? remap_cols_bytes(npz, (long long)ni * nj) : 0;
  if (ext && head + b_q + b_pt + b_col * nt > arena.bytes)
    return set_error("this step needs %zu bytes of scratch; the no-allocation scratch holds %zu "
                     "(fv3_step_dev allocates its own)", head + b_q + b_pt + b_col * nt, arena.bytes);
  char *base = ext ? arena.base : (char *)ctx_stage(c, head);
  if (!base) return 1;
  double *cmax = (double *)(base + o_cmax);
  T *frac = (T *)(base + o_frac);
  int *nspl = (int *)(base + o_ns), *flag = c.graph_safe ? c.step_flag : (int *)(base + o_flag);
  if (device_nsplit<T>(c, cx, cy, c.step_per_level, c.step_per_level ? 0 : c.step_q_split, cmax, nspl, frac))
    return 1;
  int cap = nsplit_cap;
  if (cap <= 0) {
    std::vector<int> h(npz);
    FV3_HIP(hipMemcpyAsync(h.data(), nspl, sizeof(int) * npz, hipMemcpyDeviceToHost, s));
    FV3_HIP(hipStreamSynchronize(s));
    cap = *std::max_element(h.begin(), h.end());
    if (cap < 1) return set_error("nsplit < 1 from a non-finite Courant number");
    c.last_nsplit = h;
  }
  // The FP32 form remaps out of place: the last transport pass of every
  // level lands in q_tmp (the caller's q_out serves as the alternate pass
  // buffer), and the remap writes q_out.
  T *q_tmp = nullptr, *pt_src = nullptr;
  std::vector<RemapCols> cols(nt);  // column data of every tile
  if (fast_remap) {
    if (remap_fast_check(npz, (long long)ni * nj, d.s2, kord_tr, kord_tm, pt != nullptr)) return 1;
    char *sc = ext ? arena.base + head : (char *)ctx_scratch(c, b_q + b_pt + b_col * nt);
    if (!sc) return 1;
    q_tmp = (T *)sc;
    pt_src = pt ? (T *)(sc + b_q) : nullptr;
    for (int t = 0; t < nt; t++) cols[t] = remap_cols_carve(sc + b_q + b_pt + b_col * t, npz, (long long)ni * nj);
  } else if (cap > 1) {
    q_tmp = (T *)ctx_scratch(c, sizeof(T) * qsize);
    if (!q_tmp) return 1;
  }
  FusedArgs<T> a{};
  a.is = d.is; a.ie = d.ie; a.js = d.js; a.je = d.je;
  a.isd = d.isd; a.ied = d.ied; a.jsd = d.jsd; a.jed = d.jed;
  a.npx = d.npx; a.npy = d.npy; a.nid = d.nid;
  a.s2 = d.s2; a.sx = d.sx; a.sy = d.sy; a.smx = d.smx; a.smy = d.smy;
  a.corners = d.bounded ? 0u : d.corners;
  a.edge = d.edge;
  a.npz = npz; a.nq = nq; a.hord = hord; a.cap = cap;
  a.ntiles = nt;
  a.tq = d.s2 * npz * nq;
  a.tcx = d.sx * npz;
  a.tcy = d.sy * npz;
  a.tmx = d.smx * npz;
  a.tmy = d.smy * npz;
  a.tdp = d.s2 * npz;
  a.H = std::min(nj, c.fused_h > 0 ? c.fused_h : 60);
  a.nstrip = (ni + FT_BX - 1) / FT_BX;
  a.nvblk = (nj + a.H - 1) / a.H;
  // interior strips: every lane's column i0-3..i0+60 lies in the compute
  // domain and, on a cube face, 3 or more cells inside the tile; their lean
  // rows lt..lb keep every input row j0-3..j0+H+2 in the domain and every
  // stencil row off the cube-edge rows, so the lean kernel needs no halo,
  // corner or cube-edge code
  {
    const bool edges = d.edge || (!d.bounded && d.corners != 0u);
    auto strip_ok = [&](int st) {
      const int i0 = d.is + st * FT_BX;
      const int lo = i0 - NG, hi = i0 - NG + FT_LANES - 1;
      return lo >= d.is && hi <= d.ie && (!edges || (lo >= 3 && hi <= d.npx - 3));
    };
    // (js + 6: the output row of the first input row, three above the first
    // lean row, is itself a domain row, so no row index of a lean block clamps)
    a.lt = std::max(d.js + 2 * NG, edges ? 8 : d.js + 2 * NG);
    a.lb = std::min(d.je - NG, edges ? d.npy - 4 : d.je - NG);
    a.s_lo = 1; a.s_hi = 0; a.nmid = 0;
    int s0 = -1, s1 = -2;
    for (int st = 0; st < a.nstrip; st++)
      if (strip_ok(st)) { if (s0 < 0) s0 = st; s1 = st; }
    if (s0 >= 0 && a.lb >= a.lt) {
      a.s_lo = s0;
      a.s_hi = s1;
      a.nmid = (a.lb - a.lt + a.H) / a.H;
    }
  }
  a.nsplit = nspl;
  a.q_in = q_in;
  a.q_out = fast_remap ? q_tmp : q_out;
  a.q_tmp = fast_remap ? q_out : q_tmp;
  a.cx = cx; a.cy = cy; a.mfx = mfx; a.mfy = mfy; a.dp1 = dp1;
  const GridDev<T> &g = grid_of<T>(c);
  a.area = g.area; a.rarea = g.rarea; a.dxa = g.dxa; a.dya = g.dya; a.dx = g.dx; a.dy = g.dy;
  a.sin_sg = g.sin_sg; a.ms2 = g.s2;
  a.tgs2 = g.s2;
  a.tgdx = g.sdx;
  a.tgdy = g.sdy;
  if (sizeof(T) == 4) {
    a.gxp = (const T *)c.gfac32;
    a.gxm = a.gxp + (long long)(d.nid + 1) * d.njd;
    a.gyp = a.gxm + (long long)(d.nid + 1) * d.njd;
    a.gym = a.gyp + (long long)d.nid * (d.njd + 1);
    a.gex = a.gym + (long long)d.nid * (d.njd + 1);
    a.tgf = 2 * (long long)(d.nid + 1) * d.njd + 2 * (long long)d.nid * (d.njd + 1) + (long long)d.nid * d.njd;
  }
  const bool timing = c.timing && !c.graph_safe;
  if (timing) FV3_HIP(hipEventRecord(c.ev[0], s));
  // the remap's column prep (thicknesses only) and the copies it needs run on
  // their own stream beside the transport; the field kernel joins it below
  const bool early = fast_remap && c.concurrent_kinds && c.aux[3];
  hipStream_t sp = early ? c.aux[3] : s;
  auto remap_prep = [&]() -> int {
    FV3_HIP(hipMemcpyAsync(delp_out, dp1, sizeof(double) * d.s2 * npz * nt, hipMemcpyDeviceToDevice, sp));
    if (pt) FV3_HIP(hipMemcpyAsync(pt_src, pt, sizeof(T) * d.s2 * npz * nt, hipMemcpyDeviceToDevice, sp));
    for (int t = 0; t < nt; t++)
      if (remap_fast_prep_f32(c, npz, ak, bk, ptop, dp1 + t * d.s2 * npz, delp_out + t * d.s2 * npz,
                              cols[t], sp))
        return 1;
    return 0;
  };
  if (early) {
    FV3_HIP(hipEventRecord(c.fork_ev, s));
    FV3_HIP(hipStreamWaitEvent(sp, c.fork_ev, 0));
    if (remap_prep()) return 1;
    FV3_HIP(hipEventRecord(c.prep_ev, sp));
  }
  // FP32 pipelining: the tracers go through the passes in chunks of cq, and
  // each chunk's remap runs on aux[3] (behind the column prep) while the
  // next chunk is transported; pt is remapped with the last chunk
  int cq = nq;
  if constexpr (std::is_same<T, float>::value) {
    if (early && c.pipe_chunk != 0) {
      cq = c.pipe_chunk > 0 ? c.pipe_chunk : fused32_round(c.fused_cfg);
      cq = std::max(1, std::min(cq, nq));
    }
  }
  const bool pipe = cq < nq;
  // each pass: the edge kinds fork onto auxiliary streams and join back
  // before the next pass (which reads every face's previous pass); when
  // pipelined they share two streams, so that with the remap's stream the
  // step uses four
  const bool conc = c.concurrent_kinds && c.aux[0] && c.aux[1] && c.aux[2];
  const KindStreams ks = !conc ? KindStreams{{s, s, s, s}}
                         : (pipe ? KindStreams{{s, c.aux[0], c.aux[0], c.aux[1]}}
                                 : KindStreams{{s, c.aux[0], c.aux[1], c.aux[2]}});
  const int nfork = !conc ? 0 : (pipe ? 2 : 3);
  const T *const pass_in = a.q_in;
  T *const pass_out = a.q_out, *const pass_tmp = a.q_tmp;
  const long long tq = a.tq, tl = d.s2 * npz;
  for (int q0 = 0; q0 < nq; q0 += cq) {
    const int nqc = std::min(cq, nq - q0);
    const long long qoff = (long long)q0 * npz * d.s2;
    a.nq = nqc;
    a.q_in = pass_in + qoff;
    a.q_out = pass_out + qoff;
    a.q_tmp = pass_tmp ? pass_tmp + qoff : nullptr;
    for (int it = 1; it <= cap; it++) {
      a.it = it;
      if (nfork) {
        FV3_HIP(hipEventRecord(c.fork_ev, s));
        for (int i = 0; i < nfork; i++) FV3_HIP(hipStreamWaitEvent(c.aux[i], c.fork_ev, 0));
      }
      int rc;
      if constexpr (std::is_same<T, double>::value)
        rc = launch_fused_f64(a, ks);
      else
        rc = launch_fused_f32(a, c.fused_cfg, ks);
      if (rc) return rc;
      for (int i = 0; i < nfork; i++) {
        FV3_HIP(hipEventRecord(c.join_ev[i], c.aux[i]));
        FV3_HIP(hipStreamWaitEvent(s, c.join_ev[i], 0));
      }
    }
    if constexpr (std::is_same<T, float>::value) {
      if (pipe) {
        FV3_HIP(hipEventRecord(c.chunk_ev, s));
        FV3_HIP(hipStreamWaitEvent(sp, c.chunk_ev, 0));
        const bool last = q0 + nqc >= nq;
        for (int t = 0; t < nt; t++)
          if (remap_fast_fields_f32(c, npz, nqc, q_tmp + t * tq + qoff, q_out + t * tq + qoff,
                                    last && pt_src ? pt_src + t * tl : nullptr, last && pt ? pt + t * tl : nullptr,
                                    kord_tr, kord_tm, t_min, cols[t], sp))
            return 1;
      }
    }
  }
  a.nq = nq;
  if (nsplit_cap > 0) {
    hipLaunchKernelGGL(k_nsplit_check, dim3(1), dim3(1), 0, s, nspl, npz, cap, flag);
    FV3_HIP(hipGetLastError());
  }
  if (timing) FV3_HIP(hipEventRecord(c.ev[1], s));
  // remap: source thickness is dp1, target lands in delp_out; pipelined, the
  // remap phase is the wait for the last chunks' remaps
  if (pipe) {
    if (timing) FV3_HIP(hipEventRecord(c.ev[2], s));
    FV3_HIP(hipEventRecord(c.prep_ev, sp));
    FV3_HIP(hipStreamWaitEvent(s, c.prep_ev, 0));
  } else if (fast_remap) {
    if (early)
      FV3_HIP(hipStreamWaitEvent(s, c.prep_ev, 0));
    else if (remap_prep())
      return 1;
  } else if (remap_phase) {
    FV3_HIP(hipMemcpyAsync(delp_out, dp1, sizeof(double) * d.s2 * npz * nt, hipMemcpyDeviceToDevice, s));
  }
  if (timing && !pipe) FV3_HIP(hipEventRecord(c.ev[2], s));
  if constexpr (std::is_same<T, float>::value) {
    if (fast_remap && !pipe) {
      for (int t = 0; t < nt; t++) {
        if (remap_fast_fields_f32(c, npz, nq, q_tmp + t * tq, q_out + t * tq, pt_src ? pt_src + t * tl : nullptr,
                                  pt ? pt + t * tl : nullptr, kord_tr, kord_tm, t_min, cols[t], s))
          return 1;
      }
    }
  }
  if (!fast_remap && remap_phase) {
    std::vector<int> kt(nq, kord_tr);
    if (remap_tracers_plain<T>(c, npz, nq, ak, bk, ptop, delp_out, q_out, pt, kt.data(), kord_tm,
                               t_min, 0))
      return 1;
  }
  if (timing) {
    FV3_HIP(hipEventRecord(c.ev[3], s));
    c.last_passes = cap;
    c.ev_valid = true;
  }
  if (nsplit_cap > 0 && !c.graph_safe) {
    int hf = 0;
    FV3_HIP(hipMemcpyAsync(&hf, flag, sizeof(int), hipMemcpyDeviceToHost, s));
    FV3_HIP(hipStreamSynchronize(s));
    if (hf) return set_error("a level needs more sub-cycles than nsplit_cap=%d", nsplit_cap);
  }
  return 0;
}

int remap_only_f32(Ctx &c, int nq, const float *q_src, float *q_dst, float *pt, const double *dp1,
                   double *delp_out, const double *ak, const double *bk, double ptop, int kord_tr,
                   int kord_tm, double t_min) {
  const Dims &d = c.d;
  const int npz = c.npz, nt = c.ntiles, ni = d.ie - d.is + 1, nj = d.je - d.js + 1;
  hipStream_t s = c.stream;
  if (remap_fast_check(npz, (long long)ni * nj, d.s2, kord_tr, kord_tm, pt != nullptr)) return 1;
  const size_t b_pt = pt ? ((sizeof(float) * d.s2 * npz * nt) + 255) & ~(size_t)255 : 0;
  const size_t b_col = remap_cols_bytes(npz, (long long)ni * nj);
  char *sc = (char *)ctx_scratch(c, b_pt + b_col * nt);
  if (!sc) return 1;
  float *pt_src = pt ? (float *)sc : nullptr;
  const long long tq = d.s2 * npz * nq, tl = d.s2 * npz;
  FV3_HIP(hipMemcpyAsync(delp_out, dp1, sizeof(double) * tl * nt, hipMemcpyDeviceToDevice, s));
  if (pt) FV3_HIP(hipMemcpyAsync(pt_src, pt, sizeof(float) * tl * nt, hipMemcpyDeviceToDevice, s));
  for (int t = 0; t < nt; t++) {
    const RemapCols cols = remap_cols_carve(sc + b_pt + b_col * t, npz, (long long)ni * nj);
    if (remap_fast_prep_f32(c, npz, ak, bk, ptop, dp1 + t * tl, delp_out + t * tl, cols, s)) return 1;
    if (remap_fast_fields_f32(c, npz, nq, q_src + t * tq, q_dst + t * tq, pt_src ? pt_src + t * tl : nullptr,
                              pt ? pt + t * tl : nullptr, kord_tr, kord_tm, t_min, cols, s))
      return 1;
  }
  return 0;
}

int fused_step(Ctx &c, bool fp32, int nq, const void *q_in, void *q_out, void *pt,
               const double *dp1, double *delp_out, const void *cx, const void *cy,
               const void *mfx, const void *mfy, const double *ak, const double *bk, double ptop,
               int hord, int kord_tr, int kord_tm, double t_min, int nsplit_cap,
               const StepArena &arena) {
  if (fp32)
    return fused_step_t<float>(c, nq, (const float *)q_in, (float *)q_out, (float *)pt, dp1, delp_out,
                               (const float *)cx, (const float *)cy, (const float *)mfx,
                               (const float *)mfy, ak, bk, ptop, hord, kord_tr, kord_tm, t_min,
                               nsplit_cap, arena);
  return fused_step_t<double>(c, nq, (const double *)q_in, (double *)q_out, (double *)pt, dp1, delp_out,
                              (const double *)cx, (const double *)cy, (const double *)mfx,
                              (const double *)mfy, ak, bk, ptop, hord, kord_tr, kord_tm, t_min,
                              nsplit_cap, arena);
}

}  // namespace fv3
