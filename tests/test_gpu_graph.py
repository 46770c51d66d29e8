"""The FP32 advection + remap step captured in a hipGraph (graph-safe mode):
every replay, with eager work and synchronisations between replays, must
reproduce the eager step bit for bit (the step's kernels are deterministic),
and the deferred nsplit check must report no overflow.
"""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases

pytestmark = pytest.mark.gpu


def _dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(device="cuda", dtype=dt)


def test_graph_replays_equal_eager(gpu_ready):
    case = cases.make_case("C48L16Q3jet", n=48, npz=16, nq=3, jet=True, courant=2.2, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=21, with_pt=True)
    f32, f64 = torch.float32, torch.float64
    ctx = fa.Context(case.grids, case.npz)
    q0, pt0 = _dev(case.q, f32), _dev(case.pt, f32)
    w = {k: _dev(getattr(case, k), f32) for k in ("cx", "cy", "mfx", "mfy")}
    dp1 = _dev(case.dp1, f64)
    ak, bk = _dev(case.ak, f64), _dev(case.bk, f64)
    q_in, q_out, pt = q0.clone(), torch.zeros_like(q0), pt0.clone()
    delp = torch.zeros_like(dp1)
    cap = 4
    nid = case.n + 6

    def inner(t):  # the step writes interior cells only; halo cells keep what they held
        return t.view(case.nq, case.npz, nid, nid)[:, :, 3:-3, 3:-3]

    def step(stream=None):
        fa.step_dev(ctx, fa.FORM_FUSED32, case.nq, q_in, q_out, pt, dp1, delp, w["cx"], w["cy"], w["mfx"],
                    w["mfy"], ak, bk, case.ptop, 8, 9, 10, 184.0, cap, stream)

    step()  # eager reference (also sizes the scratch)
    torch.cuda.synchronize()
    q_ref, pt_ref, delp_ref = q_out.clone(), pt.clone(), delp.clone()

    fa.set_graph_safe(ctx, True)
    pt.copy_(pt0)
    q_out.zero_()
    step()  # warm-up in graph-safe mode, outside capture
    torch.cuda.synchronize()
    assert torch.equal(q_out, q_ref) and torch.equal(pt, pt_ref)

    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        pt.copy_(pt0)
        with torch.cuda.graph(g, stream=s):
            step(torch.cuda.current_stream().cuda_stream)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    for rep in range(4):
        # eager work between replays, then restore the inputs the step reads
        junk = torch.randn(1 << 20, device="cuda").sum()
        pt.copy_(pt0)
        q_out.fill_(-1.0)
        delp.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        dq = (inner(q_out) - inner(q_ref)).abs().max().item()
        print(f"replay {rep}: max |dq| {dq:.3e}", float(junk) * 0, flush=True)
        assert torch.equal(inner(q_out), inner(q_ref)), f"replay {rep} differs"
        assert torch.equal(pt, pt_ref) and torch.equal(delp, delp_ref)
        assert fa.step_status(ctx) == 0
    fa.set_graph_safe(ctx, False)
    ctx.close()


def test_graph_safe_requires_cap(gpu_ready):
    case = cases.make_case("C24L8Q1", n=24, npz=8, nq=1, courant=0.5, seed=2, with_pt=True)
    ctx = fa.Context(case.grids, case.npz)
    fa.set_graph_safe(ctx, True)
    f32, f64 = torch.float32, torch.float64
    q = _dev(case.q, f32)
    with pytest.raises(RuntimeError, match="nsplit_cap"):
        fa.step_dev(ctx, fa.FORM_FUSED32, 1, q, torch.zeros_like(q), _dev(case.pt, f32), _dev(case.dp1, f64),
                    torch.zeros(case.dp1.size, dtype=f64, device="cuda"), _dev(case.cx, f32), _dev(case.cy, f32),
                    _dev(case.mfx, f32), _dev(case.mfy, f32), _dev(case.ak, f64), _dev(case.bk, f64),
                    case.ptop, 8, 9, 10, 184.0, 0)
    ctx.close()


def _small(case):
    f32, f64 = torch.float32, torch.float64
    d = {k: _dev(getattr(case, k), f32) for k in ("cx", "cy", "mfx", "mfy")}
    d.update(q=_dev(case.q, f32), pt=_dev(case.pt, f32), dp1=_dev(case.dp1, f64),
             ak=_dev(case.ak, f64), bk=_dev(case.bk, f64))
    d["q_out"] = torch.zeros_like(d["q"])
    d["delp"] = torch.zeros_like(d["dp1"])
    return d


def _go(ctx, case, d, form=fa.FORM_FUSED32, nq=None, cap=3):
    fa.step_dev(ctx, form, nq or case.nq, d["q"], d["q_out"], d["pt"], d["dp1"], d["delp"], d["cx"],
                d["cy"], d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, 8, 9, 10, 184.0, cap)


def test_graph_safe_arenas_do_not_move(gpu_ready):
    # a graph-safe context never frees or moves its scratch: a later call that
    # would need it bigger fails with a message instead of leaving captured
    # graphs with a freed address; the FP64 fused form (its remap synchronises
    # the host) is refused in graph-safe mode
    small = cases.make_case("C24L8Q2", n=24, npz=8, nq=2, courant=0.5, seed=3, with_pt=True)
    big = cases.make_case("C24L8Q6", n=24, npz=8, nq=6, courant=0.5, seed=3, with_pt=True)
    ctx = fa.Context(small.grids, small.npz)
    ds = _small(small)
    _go(ctx, small, ds)  # sizes the scratch for two tracers
    torch.cuda.synchronize()
    fa.set_graph_safe(ctx, True)
    _go(ctx, small, ds)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="graph-safe mode"):
        _go(ctx, big, _small(big))
    f64 = torch.float64
    d64 = {k: (v.to(f64) if v.dtype == torch.float32 else v) for k, v in ds.items()}
    with pytest.raises(RuntimeError, match="graph-safe steps need the FP32 form"):
        _go(ctx, small, d64, form=fa.FORM_FUSED64)
    fa.set_graph_safe(ctx, False)
    _go(ctx, big, _small(big))  # eager again: the scratch may grow
    torch.cuda.synchronize()
    ctx.close()


def test_python_wrappers_check_device_arguments(gpu_ready):
    case = cases.make_case("C24L8Q1", n=24, npz=8, nq=1, courant=0.5, seed=4, with_pt=True)
    ctx = fa.Context(case.grids, case.npz)
    d = _small(case)
    bad = dict(d)
    bad["q"] = d["q"].cpu()
    with pytest.raises(TypeError, match="q_in"):
        _go(ctx, case, bad)
    bad = dict(d)
    bad["q"] = d["q"].double()
    with pytest.raises(TypeError, match="dtype"):
        _go(ctx, case, bad)
    bad = dict(d)
    bad["cx"] = d["cx"][:-5]
    with pytest.raises(ValueError, match="cx"):
        _go(ctx, case, bad)
    bad = dict(d)
    bad["q_out"] = torch.zeros(2 * d["q"].numel(), device="cuda")[::2]
    with pytest.raises(ValueError, match="contiguous"):
        _go(ctx, case, bad)
    _go(ctx, case, d)  # the valid call still runs
    torch.cuda.synchronize()
    ctx.close()
