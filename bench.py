#!/usr/bin/env python3
"""Headline measurement: FV3 tracer transport + vertical remap on one MI355X.

Workload: one C768 tile (768x768, 3-cell halo), 127 levels, 30 tracers + pt.
One step = tracer_2d_1L (hord 8, per-level nsplit(k) from a jet wind
profile) followed by the tracer (kord 9) and pt (kord 10) remap of
Lagrangian_to_Eulerian, with every input already resident in HBM.  The
timed form is the fused FP32-bulk form (FP64 thickness/pressure); the plain
FP64 form is timed beside it, and the FP32 result is compared with the FP64
one.  The CPU baseline is the FP64 CPU oracle (OpenMP) on a bounded sample of the
same workload (two tracers + pt, every level, the full tile) on all host
threads, with a one-thread figure beside it.

Prints ONE JSON line.  Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-tracer updates/s C768L127x30 tracers: FP32-batched vs plain FP64 @≤1e-5; %HBM"
HBM_PEAK = 8.0e12
VALU_PEAK = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions/s: 256 CUs x 4 SIMDs, 4 cycles each


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def device_inputs(case, torch, dtype):
    """Flat device tensors in the Fortran layouts the C-ABI takes (tiles outermost)."""
    f = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to("cuda", dt)
    lev = torch.from_numpy(case.meta["lev"]).to("cuda")
    qs = []
    for tb in case.meta["q_bases"]:  # q(i,j,k,iq,tile) = base(i,j) * lev(k)
        bases = torch.stack([torch.from_numpy(np.ascontiguousarray(b.T)) for b in tb]).to("cuda", torch.float64)
        qs.append((bases[:, None, :, :] * lev[None, :, None, None]).to(dtype))
    q = torch.stack(qs) if len(qs) > 1 else qs[0]
    pt = f(case.pt, dtype) if case.pt is not None else None
    d = dict(q_in=q.reshape(-1).contiguous(), pt=pt, dp1=f(case.dp1, torch.float64),
             cx=f(case.cx, dtype), cy=f(case.cy, dtype), mfx=f(case.mfx, dtype),
             mfy=f(case.mfy, dtype), ak=f(case.ak, torch.float64), bk=f(case.bk, torch.float64))
    d["q_out"] = torch.zeros_like(d["q_in"])
    d["delp_out"] = torch.zeros_like(d["dp1"])
    return d


def nsplit_of(case):
    """nsplit(k) = int(1 + cmax(k)) of tracer_2d_1L, from the case arrays."""
    ng, n = 3, case.n
    s5 = case.grids[0].sin_sg[ng:ng + n, ng:ng + n, 4]
    out = []
    for k in range(case.npz):
        ax = np.abs(case.cx[:n, ng:ng + n, k])
        ay = np.abs(case.cy[ng:ng + n, :n, k])
        m = np.maximum(ax, ay)
        c = m.max() if (k + 1) < case.npz // 6 else (m + 1.0 - s5).max()
        out.append(int(1.0 + c))
    return np.array(out)


def run_form(fa, ctx, form, d, case, steps, warmup, cap, torch, timing=False):
    stream = torch.cuda.current_stream().cuda_stream
    t_adv, t_rem = [], []

    def one():
        fa.step_dev(ctx, form, case.nq, d["q_in"], d["q_out"], d["pt"], d["dp1"], d["delp_out"],
                    d["cx"], d["cy"], d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord,
                    case.kord_tr, case.kord_tm, 184.0, cap, stream)
        d["q_in"], d["q_out"] = d["q_out"], d["q_in"]

    for _ in range(warmup):
        one()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        one()
        if timing:
            a, r, p = fa.step_timing(ctx)
            t_adv.append(a)
            t_rem.append(r)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    return ms, t_adv, t_rem


def run_graph(fa, ctx, d, case, steps, warmup, cap, torch):
    """Capture one step per buffer direction in a hipGraph and time replays."""
    fa.set_graph_safe(ctx, True)

    def one(q_in, q_out, stream):
        fa.step_dev(ctx, fa.FORM_FUSED32, case.nq, q_in, q_out, d["pt"], d["dp1"], d["delp_out"],
                    d["cx"], d["cy"], d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord,
                    case.kord_tr, case.kord_tm, 184.0, cap, stream)

    a, b = d["q_in"], d["q_out"]
    one(a, b, None)  # graph-safe warm-up outside capture
    torch.cuda.synchronize()
    graphs = []
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for src, dst in ((a, b), (b, a)):
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g, stream=s):
                one(src, dst, torch.cuda.current_stream().cuda_stream)
            graphs.append(g)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for i in range(warmup):
        graphs[i % 2].replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        graphs[(warmup + i) % 2].replay()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    bad = fa.step_status(ctx)
    fa.set_graph_safe(ctx, False)
    return ms, bad


def profiled(tag, field, n, npz, nq, cfg, rows):
    """A per-launch figure of the transport that rocprofv3 counters measured
    for this workload and launch shape: field of the '<tag>:' line of a
    summary under profiles/ whose keys match.  (None, None) when none does."""
    import glob
    want = dict(n=str(n), npz=str(npz), nq=str(nq), cfg=str(cfg), rows=str(rows))
    for p in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(p) as f:
            for line in f:
                if not line.startswith(tag + ":"):
                    continue
                kv = dict(t.split("=", 1) for t in line.split()[1:] if "=" in t)
                if all(kv.get(k) == v for k, v in want.items()):
                    return int(float(kv[field])), os.path.relpath(p, ROOT)
    return None, None


def profiled_traffic(n, npz, nq, cfg, rows):
    """HBM bytes per transport launch: FETCH_SIZE + WRITE_SIZE of every
    k_fused_tp dispatch of one step, each scaled by the same counter's
    reading for a copy kernel of known size, divided by the step's passes."""
    return profiled("bench-traffic", "bytes_per_pass", n, npz, nq, cfg, rows)


def profiled_valu(n, npz, nq, cfg, rows):
    """VALU wave instructions per transport launch (SQ_INSTS_VALU of every
    k_fused_tp dispatch of one step divided by the step's passes)."""
    return profiled("bench-valu", "valu_per_pass", n, npz, nq, cfg, rows)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case, nq_all=2, nq_one=1, reps_all=5, reps_one=3):
    """FP64 oracle (C, OpenMP over levels in the transport and rows in the
    remap) on the full tile, every level, with pt: all host threads on nq_all
    tracers (median of reps_all steps after one warm-up), and one thread on
    nq_one tracers (median of reps_one after one warm-up)."""
    import copy
    import ctypes
    from oracle import oracle
    L = oracle.lib()
    L.fv3o_set_threads.argtypes = [ctypes.c_int]
    L.fv3o_set_threads.restype = ctypes.c_int
    bases = case.meta["q_bases"][0]

    def timed(nq, reps):
        c1 = copy.copy(case)
        c1.nq = nq
        c1.q = np.asfortranarray(np.stack([bases[i % len(bases)][:, :, None] * case.meta["lev"][None, None, :]
                                           for i in range(nq)], axis=-1))
        ts = []
        for r in range(reps + 1):
            t0 = time.perf_counter()
            adv = oracle.tracer_2d(c1, "1L")
            out = oracle.remap_tracers(c1, adv["q"], c1.dp1, c1.pt)
            if r:  # the first step is the warm-up
                ts.append(time.perf_counter() - t0)
        last[nq] = out
        return float(np.median(ts)), case.n * case.n * case.npz * nq

    last = {}

    omp_env = os.environ.get("OMP_NUM_THREADS")
    nthreads = L.fv3o_set_threads(int(omp_env) if omp_env and omp_env.isdigit() else os.cpu_count())
    t_all, c_all = timed(nq_all, reps_all)
    L.fv3o_set_threads(1)
    t_one, c_one = timed(nq_one, reps_one)
    L.fv3o_set_threads(nthreads)
    return dict(value=c_all / t_all / 1e9, unit="Gcell-tracer updates/s", cores=nthreads, kind="port",
                sample=f"FP64 oracle (C restatement, OpenMP over levels and rows), full C{case.n} tile, "
                       f"{case.npz} levels, {nq_all} tracers + pt remap; median of {reps_all} steps after "
                       f"one warm-up, {t_all:.2f} s per step on {nthreads} threads",
                one_core=dict(value=c_one / t_one / 1e9, cores=1,
                              sample=f"{nq_one} tracer + pt remap, median of {reps_one} steps after one "
                                     f"warm-up, {t_one:.2f} s per step"),
                nproc=os.cpu_count(), cpu_model=_cpu_model(), omp_num_threads=omp_env), last[nq_all]


def run_sphere(args, torch, fa, cases):
    """Six C384 faces, 127 levels, 20 tracers: the fused FP32 step against the
    fused FP64 step (bit-identical to the oracle on small spheres), global
    tracer mass over the sphere, and the FP32 step time.  Prints one JSON line."""
    t0 = time.perf_counter()
    case = cases.config(4, build_q=False)
    N, L, Q, T = case.n, case.npz, case.nq, case.ntiles
    log(f"case 6xC{N} L{L} Q{Q} built in {time.perf_counter() - t0:.1f}s")
    nid = N + 6
    area = torch.stack([torch.from_numpy(np.ascontiguousarray(g.area.T)) for g in case.grids]).cuda()
    area = area[:, 3:-3, 3:-3]

    def run(form, dtype):
        ctx = fa.Context(case.grids, L)
        d = device_inputs(case, torch, dtype)
        fa.set_timing(ctx, True)
        fa.step_dev(ctx, form, Q, d["q_in"], d["q_out"], None, d["dp1"], d["delp_out"], d["cx"], d["cy"],
                    d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord, case.kord_tr, 10, 184.0, 0)
        torch.cuda.synchronize()
        passes = fa.step_timing(ctx)[2]
        return ctx, d, passes

    def mass(q, dp, absolute=False):  # per tracer, interior cells of every face
        q = q.view(T, Q, L, nid, nid)[:, :, :, 3:-3, 3:-3].double()
        q = q.abs() if absolute else q
        dp = dp.view(T, 1, L, nid, nid)[:, :, :, 3:-3, 3:-3]
        return (q * dp * area[:, None, None]).sum(dim=(0, 2, 3, 4))

    ctx64, d64, passes = run(fa.FORM_FUSED64, torch.float64)
    m0 = mass(d64["q_in"], d64["dp1"])
    mabs = mass(d64["q_in"], d64["dp1"], absolute=True)  # tracers of either sign
    m64 = mass(d64["q_out"], d64["delp_out"])
    q64 = d64["q_out"].view(T, Q, L, nid, nid)[:, :, :, 3:-3, 3:-3].clone()
    ctx64.close()
    del d64
    torch.cuda.empty_cache()
    ctx, d, _ = run(fa.FORM_FUSED32, torch.float32)
    q32 = d["q_out"].view(T, Q, L, nid, nid)[:, :, :, 3:-3, 3:-3].double()
    rel = (torch.linalg.vector_norm((q32 - q64).transpose(0, 1).reshape(Q, -1), dim=1) /
           torch.linalg.vector_norm(q64.transpose(0, 1).reshape(Q, -1), dim=1)).max().item()
    m32 = mass(d["q_out"], d["delp_out"])
    dm64 = ((m64 - m0).abs() / mabs).max().item()
    dm32 = ((m32 - m0).abs() / mabs).max().item()
    del q32, q64
    log(f"{passes} passes; FP32 vs FP64 rel L2 max {rel:.2e}; global mass drift FP64 {dm64:.2e}, FP32 {dm32:.2e}")
    extra = {}
    if args.graph:  # the step captured once per buffer direction, replayed
        ms, bad = run_graph(fa, ctx, d, case, args.steps, args.warmup, passes, torch)
        if bad:
            raise SystemExit(f"graph replays: nsplit above the captured {passes} passes")
        extra = {"graph": True}
        log(f"fused FP32 sphere, hipGraph replays: {ms:.2f} ms/step")
    else:
        ms, t_adv, t_rem = run_form(fa, ctx, fa.FORM_FUSED32, d, case, args.steps, args.warmup, passes, torch,
                                    timing=True)
        extra = {"transport_ms": round(float(np.mean(t_adv)), 3), "remap_ms": round(float(np.mean(t_rem)), 3)}
        log(f"fused FP32 sphere: {ms:.2f} ms/step (transport {np.mean(t_adv):.2f}, remap {np.mean(t_rem):.2f})")
    value = T * N * N * L * Q / (ms * 1e-3) / 1e9
    print(json.dumps({
        "metric": f"Gcell-tracer updates/s 6xC{N}L{L}x{Q} tracers, fused FP32 step", "value": round(value, 3),
        "unit": "Gcell-tracer updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"6 x C{N} L{L} x {Q} tracers, solid-body rotation over the cube corners: "
                               f"tracer_2d_1L hord 8 + kord 9 tracer remap, fused FP32 form",
                   "n": N, "npz": L, "nq": Q, "ntiles": T, "passes": passes},
        **extra, "fp32_rel_l2_max": rel, "mass_drift_fp64": dm64, "mass_drift_fp32": dm32}), flush=True)
    ctx.close()


def run_sphere_adjoint(args, torch, fa, cases):
    """The sphere workload's final step differentiated: six C384 faces, 127
    levels, tracer 0 only, FP64 forward (tracer_2d_1L + kord 9 remap) with
    its tangent-linear, then the adjoint, and the dot-product test between
    them.  Prints one JSON line."""
    t0 = time.perf_counter()
    case = cases.config(4, build_q=False)
    N, L, T = case.n, case.npz, case.ntiles
    log(f"case 6xC{N} L{L} (tracer 0) built in {time.perf_counter() - t0:.1f}s")
    d = device_inputs(case, torch, torch.float64)
    nid = N + 6
    q0 = d["q_in"].view(T, case.nq, L, nid, nid)[:, :1].contiguous().view(-1)
    del d["q_in"], d["q_out"]
    ctx = fa.Context(case.grids, L)
    g = torch.Generator(device="cuda").manual_seed(7)

    def field():
        f = torch.randn(T, 1, L, nid, nid, device="cuda", dtype=torch.float64, generator=g)
        f[..., :3, :] = 0
        f[..., -3:, :] = 0
        f[..., :, :3] = 0
        f[..., :, -3:] = 0
        return f.view(-1)

    dq, y = field(), field()
    q1, dq1, qb = torch.empty_like(q0), torch.empty_like(q0), torch.empty_like(q0)
    delp = torch.empty_like(d["dp1"])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fa.step_tl_dev(ctx, True, 0, 1, q0, dq, q1, dq1, d["dp1"], delp, d["cx"], d["cy"], d["mfx"], d["mfy"],
                   d["ak"], d["bk"], case.ptop, case.hord, 9)
    torch.cuda.synchronize()
    t_tl = time.perf_counter() - t0
    t0 = time.perf_counter()
    fa.step_ad_dev(ctx, True, 0, 1, q0, y, qb, d["dp1"], d["cx"], d["cy"], d["mfx"], d["mfy"], d["ak"], d["bk"],
                   case.ptop, case.hord, 9)
    torch.cuda.synchronize()
    t_ad = time.perf_counter() - t0
    lhs, rhs = float(torch.dot(dq1, y)), float(torch.dot(dq, qb))
    rel = abs(lhs - rhs) / abs(lhs)
    log(f"TL {t_tl:.2f} s, AD {t_ad:.2f} s, <TL dx, y> {lhs:.15e} <dx, AD y> {rhs:.15e} rel {rel:.2e}")
    ctx.close()
    print(json.dumps({"metric": "sphere final-step adjoint", "workload": f"6xC{N} L{L}, tracer 0, FP64, "
                      "tracer_2d_1L + kord 9 remap", "tl_s": round(t_tl, 3), "ad_s": round(t_ad, 3),
                      "dot_product_rel": rel}), flush=True)


def run_mixed_precision(args, torch, fa, cases):
    """One C384 tile, 64 levels, 10 tracers, deformational winds, forced
    nsplit = 2 (tracer_2d rule, q_split 2), kord 9 remap: the fused FP32 step
    against the fused FP64 step and the FP64 CPU oracle, per tracer (relative
    L2 and Linf, mass change), and the FP32 and FP64 step times.  Prints one
    JSON line."""
    from oracle import oracle
    case = cases.config(1)
    N, L, Q = case.n, case.npz, case.nq
    nid = N + 6
    log(f"case C{N} L{L} Q{Q} q_split {case.q_split}")
    area = torch.from_numpy(np.ascontiguousarray(case.grids[0].area.T)).cuda()[3:-3, 3:-3]

    def run(form, dtype):
        ctx = fa.Context(case.grids, L)
        fa.set_transport(ctx, False, case.q_split)
        fa.set_timing(ctx, True)
        d = device_inputs(case, torch, dtype)
        q0 = d["q_in"].clone()
        fa.step_dev(ctx, form, Q, d["q_in"], d["q_out"], None, d["dp1"], d["delp_out"], d["cx"], d["cy"],
                    d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord, case.kord_tr, 10, 184.0, 0)
        torch.cuda.synchronize()
        out = d["q_out"].view(Q, L, nid, nid)[:, :, 3:-3, 3:-3].double().clone()
        dpo = d["delp_out"].view(L, nid, nid)[:, 3:-3, 3:-3].clone()
        ms, _, _ = run_form(fa, ctx, form, d, case, args.steps, args.warmup, case.q_split, torch)
        ctx.close()
        return out, dpo, q0.view(Q, L, nid, nid)[:, :, 3:-3, 3:-3].double(), ms

    q64, dp64, q0, ms64 = run(fa.FORM_FUSED64, torch.float64)
    q32, dp32, _, ms32 = run(fa.FORM_FUSED32, torch.float32)
    t0 = time.perf_counter()
    adv = oracle.tracer_2d(case, "2d")
    qo, dpo_, _ = oracle.remap_tracers(case, adv["q"], case.dp1, None)
    log(f"CPU oracle step in {time.perf_counter() - t0:.1f} s")
    qo = torch.from_numpy(np.ascontiguousarray(np.transpose(qo[3:-3, 3:-3], (3, 2, 1, 0)))).cuda()
    dp1 = torch.from_numpy(np.ascontiguousarray(case.dp1.T)).cuda()[:, 3:-3, 3:-3]
    m0 = (q0 * dp1 * area).sum(dim=(1, 2, 3))
    mabs = (q0.abs() * dp1 * area).sum(dim=(1, 2, 3))

    def errs(q, ref):
        l2 = (torch.linalg.vector_norm((q - ref).reshape(Q, -1), dim=1) /
              torch.linalg.vector_norm(ref.reshape(Q, -1), dim=1))
        li = (q - ref).abs().amax(dim=(1, 2, 3)) / ref.abs().amax(dim=(1, 2, 3))
        return [float(x) for x in l2], [float(x) for x in li]

    l2_64, li_64 = errs(q64, qo)
    l2_32, li_32 = errs(q32, qo)
    # one open tile: its mass changes by what crosses the tile edges, so the
    # FP32 form's conservation error is its mass against the FP64 form's
    m32 = (q32 * dp32 * area).sum(dim=(1, 2, 3))
    m64 = (q64 * dp64 * area).sum(dim=(1, 2, 3))
    drift32 = [float(x) for x in (m32 - m64).abs() / mabs]
    flux64 = [float(x) for x in (m64 - m0) / mabs]
    for iq in range(Q):
        log(f"tracer {iq}: FP32 rel L2 {l2_32[iq]:.2e} Linf {li_32[iq]:.2e} mass vs FP64 {drift32[iq]:.1e}; "
            f"FP64 rel L2 {l2_64[iq]:.1e}")
    cells = N * N * L * Q
    print(json.dumps({
        "metric": f"Gcell-tracer updates/s C{N}L{L}x{Q} tracers, fused FP32 vs FP64 error report",
        "value": round(cells / (ms32 * 1e-3) / 1e9, 3), "unit": "Gcell-tracer updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms32, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"C{N} tile L{L} x {Q} tracers, deformational winds, tracer_2d with "
                               f"q_split {case.q_split}, kord 9 tracer remap", "n": N, "npz": L, "nq": Q},
        "fp64_ms_per_step": round(ms64, 3), "fp32_rel_l2_vs_oracle": l2_32, "fp32_linf_vs_oracle": li_32,
        "fp64_rel_l2_vs_oracle": l2_64, "fp32_mass_vs_fp64": drift32, "fp64_tile_edge_mass_flux": flux64}),
        flush=True)


def run_next_rows(args, torch, fa, case, d):
    """The rows either side of the path at the workload's size, timed with
    events on the context stream: dyn_core's accumulation of the transport
    inputs (two acoustic sub-steps, FP32, winds derived from the workload's
    Courant numbers) and, with --io, the tracer restart and history files."""
    N, L = case.n, case.npz
    nd = N + 6
    out = {}
    ctx = fa.Context(case.grids, L)
    g = case.grids[0]
    dxa = torch.from_numpy(np.ascontiguousarray(g.dxa.T)).cuda().float()
    dya = torch.from_numpy(np.ascontiguousarray(g.dya.T)).cuda().float()
    cx = d["cx"].view(L, nd, N + 1)
    cy = d["cy"].view(L, N + 1, nd)
    u0 = torch.where(cx > 0, cx * dxa[:, 2:N + 3], cx * dxa[:, 3:N + 4]) / case.dt
    v0 = torch.where(cy > 0, cy * dya[2:N + 3, :], cy * dya[3:N + 4, :]) / case.dt
    n_split = 2
    uc = torch.stack([u0, 1.1 * u0]).contiguous()
    vc = torch.stack([v0, 0.9 * v0]).contiguous()
    del u0, v0
    delp = d["dp1"].float()
    dev = {k: torch.empty(n, dtype=torch.float32, device="cuda")
           for k, n in (("dp1", delp.numel()), ("cx", cx.numel()), ("cy", cy.numel()),
                        ("mfx", L * N * (N + 1)), ("mfy", L * N * (N + 1)))}
    bdt = case.dt / n_split
    d0 = delp.clone()
    run = lambda: fa.dyn_core_fluxes(ctx, fa.FORM_FUSED32, n_split, bdt, case.hord, uc, vc, delp, dev["dp1"],
                                     dev["cx"], dev["cy"], dev["mfx"], dev["mfy"])
    run()
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        delp.copy_(d0)
        run()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps / n_split
    # per cell and sub-step: uc, vc read; delp, cx, cy, mfx, mfy read and written
    b = N * N * L * 4 * (2 + 2 * 5)
    out["dyn_core"] = dict(ms_per_substep=round(ms, 3), achieved_GBps=round(b / ms / 1e6, 1),
                           frac_hbm=round(b / (ms * 1e-3) / HBM_PEAK, 4),
                           workload=f"C{N} L{L} FP32, {n_split} sub-steps, one launch per sub-step")
    log(f"dyn_core fluxes: {ms:.2f} ms per acoustic sub-step")
    del uc, vc, dev, d0
    if args.io:
        import tempfile
        with tempfile.TemporaryDirectory(dir="/tmp") as td:
            q = d["q_in"][: 2 * L * nd * nd]
            t0 = time.perf_counter()
            fa.tracer_restart_write(ctx, td + "/fv_tracer.res", ["sphum", "liq_wat"], q)
            tw = time.perf_counter() - t0
            back = torch.empty_like(q)
            t0 = time.perf_counter()
            fa.tracer_restart_read(ctx, td + "/fv_tracer.res", ["sphum", "liq_wat"], back)
            tr = time.perf_counter() - t0
            fb = 2 * N * N * L * 8
            out["restart_io"] = dict(tracers=2, file_bytes=fb, write_s=round(tw, 3), read_s=round(tr, 3),
                                     write_GBps=round(fb / tw / 1e9, 2), read_GBps=round(fb / tr / 1e9, 2))
            log(f"restart write {tw:.2f} s, read {tr:.2f} s for {fb / 1e9:.2f} GB")
            del back
    ctx.close()
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=768)
    ap.add_argument("--npz", type=int, default=127)
    ap.add_argument("--nq", type=int, default=30)
    ap.add_argument("--plain-steps", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calib", action="store_true",
                    help="first run a 1 GiB copy kernel (reference for HBM byte counters)")
    ap.add_argument("--launch", type=int, default=None,
                    help="fused launch configuration (default: the library's, 9); -1 tunes over all "
                         "of them and the rows per block first")
    ap.add_argument("--rows", type=int, default=0, help="rows per block (0: the library default, 60)")
    ap.add_argument("--remap", type=int, default=0, help="remap kernel (0: the library default)")
    ap.add_argument("--serial-kinds", action="store_true",
                    help="run the transport's edge kernels after the lean kernel on one stream")
    ap.add_argument("--graph", action="store_true",
                    help="time the step as hipGraph replays (graph-safe mode): the default")
    ap.add_argument("--eager", action="store_true",
                    help="time eager steps only, not hipGraph replays of the captured step")
    ap.add_argument("--sphere", action="store_true",
                    help="the six-face C384 L127 Q20 sphere instead of the C768 tile")
    ap.add_argument("--pipeline", type=int, default=None,
                    help="tracers per transport chunk, each chunk remapped beside the next chunk's "
                         "transport (-1: one round of the launch shape; 0: off, the library's default)")
    ap.add_argument("--no-next", action="store_true", help="skip the dyn_core input timing")
    ap.add_argument("--io", action="store_true", help="also time the tracer restart write and read")
    ap.add_argument("--sphere-adjoint", action="store_true",
                    help="differentiate the sphere workload's final step (tracer 0, FP64 TL + AD)")
    ap.add_argument("--mixed", action="store_true",
                    help="the C384 L64 Q10 FP32-against-FP64 error report instead of the C768 tile")
    args = ap.parse_args()
    args.graph = not args.eager  # the workload runs as a captured step under hipGraph
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} requested); "
              "across GPUs the path would shard one cube tile per GPU, which is not built",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import fv3atm_amd as fa
    from fv3atm_amd import cases

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    if args.sphere:
        run_sphere(args, torch, fa, cases)
        return
    if args.mixed:
        run_mixed_precision(args, torch, fa, cases)
        return
    if args.sphere_adjoint:
        run_sphere_adjoint(args, torch, fa, cases)
        return
    if args.calib:  # reads 1 GiB and writes 1 GiB in one elementwise kernel
        x = torch.ones(1 << 28, dtype=torch.int32, device="cuda")
        y = torch.empty_like(x)
        torch.bitwise_not(x, out=y)
        torch.cuda.synchronize()
        del x, y
        log("calibration kernel: 1 GiB read + 1 GiB write (int32 bitwise_not)")
    t0 = time.perf_counter()
    case = cases.config(2, n=args.n, npz=args.npz, nq=args.nq, build_q=False)
    nspl = nsplit_of(case)
    S = int(nspl.sum())
    N, L, Q = case.n, case.npz, case.nq
    log(f"case C{N} L{L} Q{Q} built in {time.perf_counter() - t0:.1f}s; nsplit histogram "
        f"{dict(zip(*np.unique(nspl, return_counts=True)))} S={S}")
    cap = int(nspl.max())
    out = {}
    keep32 = None

    # ---- plain FP64 form (timing + reference for the FP32 error)
    ref64 = None
    if not args.no_plain:
        ctx64 = fa.Context(case.grids, L)
        d64 = device_inputs(case, torch, torch.float64)
        q0_64 = d64["q_in"].clone()
        pt0 = d64["pt"].clone()
        fa.step_dev(ctx64, fa.FORM_PLAIN64, Q, d64["q_in"], d64["q_out"], d64["pt"], d64["dp1"],
                    d64["delp_out"], d64["cx"], d64["cy"], d64["mfx"], d64["mfy"], d64["ak"],
                    d64["bk"], case.ptop, case.hord, case.kord_tr, case.kord_tm, 184.0, cap)
        torch.cuda.synchronize()
        ref64 = dict(q=d64["q_out"].clone(), pt=d64["pt"].clone(), delp=d64["delp_out"].clone(),
                     q0=q0_64)
        d64["pt"].copy_(pt0)
        del pt0
        ms64, _, _ = run_form(fa, ctx64, fa.FORM_PLAIN64, d64, case, args.plain_steps, 0, cap, torch)
        log(f"plain FP64: {ms64:.2f} ms/step")
        out["plain64_ms_per_step"] = round(ms64, 3)
        # the fused FP64 form (the FP32 form's kernels in FP64, per-column remap)
        ms64f, _, _ = run_form(fa, ctx64, fa.FORM_FUSED64, d64, case, args.plain_steps, 1, cap, torch)
        log(f"fused FP64: {ms64f:.2f} ms/step")
        out["fused64_ms_per_step"] = round(ms64f, 3)
        del d64
        ctx64.close()
        torch.cuda.empty_cache()

    # ---- fused FP32 form
    ctx = fa.Context(case.grids, L)
    d = device_inputs(case, torch, torch.float32)
    if ref64 is not None:  # one step from the same state, compared with FP64
        fa.step_dev(ctx, fa.FORM_FUSED32, Q, d["q_in"], d["q_out"], d["pt"], d["dp1"], d["delp_out"],
                    d["cx"], d["cy"], d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord,
                    case.kord_tr, case.kord_tm, 184.0, cap)
        torch.cuda.synchronize()
        nid = N + 6
        shp = (Q, L, nid, nid)
        g32 = d["q_out"].view(shp)[:, :, 3:-3, 3:-3].double()
        g64 = ref64["q"].view(shp)[:, :, 3:-3, 3:-3]
        num = torch.linalg.vector_norm((g32 - g64).reshape(Q, -1), dim=1)
        den = torch.linalg.vector_norm(g64.reshape(Q, -1), dim=1)
        rel = (num / den).cpu().numpy()
        linf = ((g32 - g64).abs().amax(dim=(1, 2, 3)) / g64.abs().amax(dim=(1, 2, 3))).cpu().numpy()
        area = torch.from_numpy(np.ascontiguousarray(case.grids[0].area.T)).cuda()[3:-3, 3:-3]
        delp = ref64["delp"].view(L, nid, nid)[:, 3:-3, 3:-3]
        dp1 = d["dp1"].view(L, nid, nid)[:, 3:-3, 3:-3]
        q0 = ref64["q0"].view(shp)[:, :, 3:-3, 3:-3]
        m0 = (q0 * dp1 * area).sum(dim=(1, 2, 3))
        dm64 = (g64 * delp * area).sum(dim=(1, 2, 3)) - m0
        dm32 = (g32 * delp * area).sum(dim=(1, 2, 3)) - m0
        mabs = (q0.abs() * dp1 * area).sum(dim=(1, 2, 3))
        mass_err = ((dm32 - dm64).abs() / mabs).max().item()
        out["fp32_rel_l2_max"] = float(rel.max())
        out["fp32_linf_rel_max"] = float(linf.max())
        out["fp32_mass_err_per_step"] = float(mass_err)
        log(f"FP32 vs FP64: rel L2 max {rel.max():.2e}, Linf rel max {linf.max():.2e}, "
            f"mass error {mass_err:.2e}")
        keep32 = dict(q=d["q_out"].view(shp)[:2].double().cpu().numpy(), pt=d["pt"].view(L, nid, nid).double().cpu().numpy())
        del ref64, g32, g64
        torch.cuda.empty_cache()
        d = device_inputs(case, torch, torch.float32)
    fa.set_timing(ctx, True)
    launch, rows = (fa.DEFAULT_LAUNCH_CFG if args.launch is None else args.launch), args.rows
    if launch < 0:  # untimed tuning of the launch shape, then rows per block
        best = None
        for cfg in range(fa.N_LAUNCH_CFG):
            fa.set_launch(ctx, cfg, rows, args.remap)
            _, ta_c, _ = run_form(fa, ctx, fa.FORM_FUSED32, d, case, 2, 1, cap, torch, timing=True)
            t = float(np.min(ta_c))
            log(f"launch cfg {cfg} rows {rows or 'default'}: transport {t:.2f} ms")
            if best is None or t < best[0]:
                best = (t, cfg, rows)
        for r in (42, 48, 60, 72, 76, 90, 95, 127):
            if r == rows or r > N:
                continue
            fa.set_launch(ctx, best[1], r, args.remap)
            _, ta_c, _ = run_form(fa, ctx, fa.FORM_FUSED32, d, case, 2, 1, cap, torch, timing=True)
            t = float(np.min(ta_c))
            log(f"launch cfg {best[1]} rows {r}: transport {t:.2f} ms")
            if t < best[0]:
                best = (t, best[1], r)
        launch, rows = best[1], best[2]
    fa.set_launch(ctx, launch, rows, args.remap)
    fa.set_concurrency(ctx, not args.serial_kinds)
    if args.pipeline is not None:
        fa.set_pipeline(ctx, args.pipeline)
    out["launch_cfg"] = launch
    out["concurrent_kinds"] = not args.serial_kinds
    out["remap_kernel"] = args.remap
    out["rows_per_block"] = rows or 60
    ms, t_adv, t_rem = run_form(fa, ctx, fa.FORM_FUSED32, d, case, args.steps, args.warmup, cap,
                                torch, timing=True)
    log(f"fused FP32: {ms:.2f} ms/step (transport {np.mean(t_adv):.2f} ms, remap {np.mean(t_rem):.2f} ms)")
    if args.graph:
        ms_eager = ms
        ms, bad = run_graph(fa, ctx, d, case, args.steps, args.warmup, cap, torch)
        assert bad == 0, "a level needed more sub-cycles than the captured nsplit cap"
        log(f"fused FP32 as hipGraph replays: {ms:.2f} ms/step (eager {ms_eager:.2f})")
        out["eager_ms_per_step"] = round(ms_eager, 3)
        out["graph"] = True

    cells = N * N * L * Q
    value = cells / (ms * 1e-3) / 1e9
    w = 4
    b_adv = S * N * N * (2 * w * Q + 24)
    b_rem = N * N * ((Q + 1) * 2 * w * L + 16 * (L + 1))
    ta = float(np.mean(t_adv)) * 1e-3
    # one launch = one sub-cycle pass over every level (the interior and frame
    # grids of k_fused_tp back to back); cap passes per step
    traffic, tsrc = profiled_traffic(N, L, Q, launch, rows or 60)
    roof = dict(bound="hbm", achieved=round(b_adv / ta / 1e9, 1), peak=HBM_PEAK / 1e9,
                unit="GB/s", frac=round(b_adv / ta / HBM_PEAK, 4), traffic=traffic)
    out["algorithmic_bytes_per_launch"] = b_adv // cap
    out["launches_per_step"] = cap
    if tsrc:
        out["traffic_source"] = tsrc
    # the bound that holds: VALU issue (wave64 instructions of the launch's
    # kernels, from the committed counter summary, over the live launch time)
    valu, vsrc = profiled_valu(N, L, Q, launch, rows or 60)
    if valu:
        rate = valu * cap / ta
        out["valu_issue"] = dict(achieved=round(rate / 1e9, 1), peak=round(VALU_PEAK / 1e9, 1),
                                 unit="G wave-instr/s", frac=round(rate / VALU_PEAK, 4),
                                 instr_per_launch=valu, source=vsrc)
    out_line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-tracer updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"C{N} tile L{L} x {Q} tracers + pt: tracer_2d_1L (hord 8, "
                               f"per-level nsplit, S={S}) + Lagrangian_to_Eulerian tracer/pt remap "
                               f"(kord 9/10), fused FP32 form",
                   "n": N, "npz": L, "nq": Q, "nsplit_sum": S, "nsplit_max": cap,
                   "form": "fused32"},
        "roofline": roof,
        "transport_ms": round(ta * 1e3, 3), "remap_ms": round(float(np.mean(t_rem)), 3),
        "hbm_frac_step": round((b_adv + b_rem) / (ms * 1e-3) / HBM_PEAK, 4),
        "algorithmic_bytes_per_step": b_adv + b_rem,
    }
    out_line.update(out)
    if "plain64_ms_per_step" in out:
        out_line["speedup_vs_plain64"] = round(out["plain64_ms_per_step"] / ms, 3)
        out_line["speedup_vs_fused64"] = round(out["fused64_ms_per_step"] / ms, 3)
    ctx.close()
    torch.cuda.empty_cache()
    if not args.no_next:
        out_line.update(run_next_rows(args, torch, fa, case, d))
    del d
    torch.cuda.empty_cache()
    if not args.no_cpu:
        cb, ref = cpu_baseline(case)
        log(f"cpu baseline {cb['value']:.4f} Gcell-tracer updates/s")
        out_line["cpu_baseline"] = cb
        if keep32 is not None:
            # the FP32 step of the workload against the FP64 oracle itself, full
            # tile and every level, for the tracers the oracle leg ran (and pt)
            qo, _, pto = ref
            errs = []
            for iq in range(qo.shape[3]):
                a = keep32["q"][iq][:, 3:-3, 3:-3]                      # (k, j, i)
                b = np.transpose(qo[3:-3, 3:-3, :, iq], (2, 1, 0))
                errs.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
            a = keep32["pt"][:, 3:-3, 3:-3]
            b = np.transpose(pto[3:-3, 3:-3, :], (2, 1, 0))
            e_pt = float(np.linalg.norm(a - b) / np.linalg.norm(b))
            out_line["fp32_vs_oracle"] = dict(tracers=len(errs), rel_l2=errs, pt_rel_l2=e_pt,
                                              sample=f"full C{N} tile, {L} levels, tracers 0..{len(errs) - 1} and pt")
            log(f"FP32 vs FP64 oracle at full size: rel L2 {max(errs):.2e}, pt {e_pt:.2e}")
    else:
        out_line["cpu_baseline"] = None
    print(json.dumps(out_line), flush=True)


if __name__ == "__main__":
    main()
