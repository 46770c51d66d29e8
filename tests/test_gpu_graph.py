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
