"""GPU tests of two benchmark shapes at sizes the oracle runs quickly:

* the plain-vs-mixed-precision shape (one tile, 10 distinct tracers: cosine
  bells, slotted cylinders and smooth fields; deformational winds at Courant
  0.8 with tracer_2d's forced q_split = 2; kord 9 remap), reduced to C96 L16:
  the FP32 form against the FP64 oracle per tracer (relative L2 <= 1e-5,
  Linf reported) and its tracer mass against the FP64 form's;
* the full-sphere shape (six faces, solid-body rotation over the cube
  corners) reduced to 6 x C48 L8 Q4, captured in a hipGraph: every replay
  equal to the eager step bit for bit.
"""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases
from oracle import oracle
from tests._util import assert_close, interior, rel_l2

pytestmark = pytest.mark.gpu
NG = 3


def _dev(a, dt):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(device="cuda", dtype=dt)


def _step(case, form, per_level, q_split):
    dt = torch.float32 if form == fa.FORM_FUSED32 else torch.float64
    ctx = fa.Context(case.grids, case.npz)
    fa.set_transport(ctx, per_level, q_split)
    q_in = _dev(case.q, dt)
    q_out = torch.zeros_like(q_in)
    dp1 = _dev(case.dp1, torch.float64)
    delp = torch.zeros_like(dp1)
    w = [_dev(getattr(case, k), dt) for k in ("cx", "cy", "mfx", "mfy")]
    fa.step_dev(ctx, form, case.nq, q_in, q_out, None, dp1, delp, *w, _dev(case.ak, torch.float64),
                _dev(case.bk, torch.float64), case.ptop, case.hord, case.kord_tr, case.kord_tm, 184.0, 0)
    torch.cuda.synchronize()
    ctx.close()
    return (q_out.double().cpu().numpy().reshape(case.q.shape, order="F"),
            delp.cpu().numpy().reshape(case.dp1.shape, order="F"))


def test_mixed_precision_deformational_c96(gpu_ready):
    case = cases.make_case("C96L16Q10deform", n=96, npz=16, nq=10, wind="deform", courant=0.8,
                           q_split=2, delp_kind=1, delp_amp=(0.05, 0.1), seed=5)
    print("[configs] C96 L16 Q10 deformational, q_split 2", flush=True)
    adv = oracle.tracer_2d(case, "2d")
    assert (adv["nsplit"] == 2).all()
    q_ref, delp_ref, _ = oracle.remap_tracers(case, adv["q"], case.dp1)
    q32, delp32 = _step(case, fa.FORM_FUSED32, False, 2)
    q64, _ = _step(case, fa.FORM_FUSED64, False, 2)
    np.testing.assert_allclose(interior(delp32), interior(delp_ref), rtol=1e-14)
    area = case.grids[0].area[NG:-NG, NG:-NG, None]
    dp = interior(delp_ref)
    for iq in range(case.nq):
        r = interior(q_ref[..., iq])
        e = rel_l2(interior(q32[..., iq]), r)
        linf = np.abs(interior(q32[..., iq]) - r).max() / np.abs(r).max()
        m32 = (interior(q32[..., iq]) * dp * area).sum()
        m64 = (interior(q64[..., iq]) * dp * area).sum()
        mabs = (np.abs(interior(q64[..., iq])) * dp * area).sum()
        print(f"tracer {iq}: rel L2 {e:.2e}, Linf {linf:.2e}, mass vs FP64 {abs(m32 - m64) / mabs:.2e}",
              flush=True)
        assert e <= 1e-5, (iq, e)
        assert abs(m32 - m64) <= 1e-6 * mabs, iq
        assert_close(interior(q64[..., iq]), interior(q_ref[..., iq]), 1e-12, f"FP64 tracer {iq}",
                     floor=1e-300)


def test_six_face_graph_replays_equal_eager(gpu_ready):
    case = cases.make_case("6xC48L8Q4", n=48, npz=8, nq=4, ntiles=6, courant=1.5, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=7, with_pt=True)
    print("[configs] 6 x C48 L8 Q4 sphere under hipGraph", flush=True)
    f32, f64 = torch.float32, torch.float64
    ctx = fa.Context(case.grids, case.npz)
    q0, pt0 = _dev(case.q, f32), _dev(case.pt, f32)
    w = [_dev(getattr(case, k), f32) for k in ("cx", "cy", "mfx", "mfy")]
    dp1 = _dev(case.dp1, f64)
    ak, bk = _dev(case.ak, f64), _dev(case.bk, f64)
    q_in, q_out, pt, delp = q0.clone(), torch.zeros_like(q0), pt0.clone(), torch.zeros_like(dp1)
    nid = case.n + 6

    def inner(t):  # tiles outermost; interior cells only
        return t.view(6, case.nq, case.npz, nid, nid)[..., NG:-NG, NG:-NG]

    def step(stream=None):
        fa.step_dev(ctx, fa.FORM_FUSED32, case.nq, q_in, q_out, pt, dp1, delp, *w, ak, bk, case.ptop,
                    8, 9, 10, 184.0, 4, stream)

    step()
    torch.cuda.synchronize()
    q_ref, pt_ref, delp_ref = q_out.clone(), pt.clone(), delp.clone()
    fa.set_graph_safe(ctx, True)
    pt.copy_(pt0)
    step()  # graph-safe warm-up outside capture
    torch.cuda.synchronize()
    assert torch.equal(inner(q_out), inner(q_ref))
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
        junk = torch.randn(1 << 20, device="cuda").sum()
        pt.copy_(pt0)
        q_out.fill_(-1.0)
        delp.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        print(f"replay {rep}", float(junk) * 0, flush=True)
        assert torch.equal(inner(q_out), inner(q_ref)), f"replay {rep} differs"
        assert torch.equal(pt, pt_ref) and torch.equal(delp, delp_ref)
        assert fa.step_status(ctx) == 0
    fa.set_graph_safe(ctx, False)
    ctx.close()
