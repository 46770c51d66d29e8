"""Results must not depend on memory the step does not own or must not read.

- The context's scratch and staging arenas are left full of another case's
  data, and q_out (the step's output buffer) is filled with NaN before the
  step: the outputs equal a fresh context's bit for bit.
- On six faces the fused and plain forms fill every halo from the
  neighbouring faces: q_in's halo strips and corner blocks set to NaN change
  nothing in the outputs.
"""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases

pytestmark = pytest.mark.gpu
NG = 3


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(device="cuda", dtype=dtype)


def _inputs(case, dt):
    d = dict(q_in=_dev(case.q, dt), pt=_dev(case.pt, dt), dp1=_dev(case.dp1, torch.float64))
    d.update({k: _dev(getattr(case, k), dt) for k in ("cx", "cy", "mfx", "mfy")})
    d["ak"], d["bk"] = _dev(case.ak, torch.float64), _dev(case.bk, torch.float64)
    d["q_out"] = torch.zeros_like(d["q_in"])
    d["delp_out"] = torch.zeros_like(d["dp1"])
    return d


def _step(ctx, case, form, d):
    fa.step_dev(ctx, form, case.nq, d["q_in"], d["q_out"], d["pt"], d["dp1"], d["delp_out"], d["cx"], d["cy"],
                d["mfx"], d["mfy"], d["ak"], d["bk"], case.ptop, case.hord, case.kord_tr, case.kord_tm, 184.0, 0)
    torch.cuda.synchronize()


def _interior_q(case, t):
    nd = case.n + 2 * NG
    shp = (nd, nd, case.npz, case.nq) + ((case.ntiles,) if case.ntiles > 1 else ())
    return t.double().cpu().numpy().reshape(shp, order="F")[NG:-NG, NG:-NG]


@pytest.mark.parametrize("form", [fa.FORM_FUSED32, fa.FORM_FUSED64])
def test_stale_scratch_and_poisoned_output(gpu_ready, form):
    dt = torch.float32 if form == fa.FORM_FUSED32 else torch.float64
    case = cases.make_case("C40L10Q3jet", n=40, npz=10, nq=3, jet=True, courant=2.4, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=21, with_pt=True)
    other = cases.make_case("C40L10Q3", n=40, npz=10, nq=3, courant=0.9, delp_kind=1, delp_amp=(0.1, 0.3),
                            seed=22, with_pt=True)
    print(f"[poison] stale scratch, form {form}", flush=True)
    fresh = fa.Context(case.grids, case.npz)
    d0 = _inputs(case, dt)
    _step(fresh, case, form, d0)
    fresh.close()
    ctx = fa.Context(other.grids, other.npz)
    # leave the arenas holding another case's state (and a host-array call's staging)
    _step(ctx, other, form, _inputs(other, dt))
    a = {k: np.array(getattr(other, k), order="F", copy=True) for k in ("q", "dp1", "cx", "cy", "mfx", "mfy")}
    fa.tracer_2d_1L(ctx, a["q"], a["dp1"], a["mfx"], a["mfy"], a["cx"], a["cy"], other.n + 1, other.n + 1,
                    other.npz, other.nq, other.hord)
    ctx.close()
    ctx = fa.Context(case.grids, case.npz)
    _step(ctx, other, form, _inputs(other, dt))
    d1 = _inputs(case, dt)
    d1["q_out"].fill_(float("nan"))
    d1["delp_out"].fill_(float("nan"))
    _step(ctx, case, form, d1)
    ctx.close()
    np.testing.assert_array_equal(_interior_q(case, d1["q_out"]), _interior_q(case, d0["q_out"]))
    assert torch.equal(d1["pt"], d0["pt"])
    nd = case.n + 2 * NG
    dl1 = d1["delp_out"].cpu().numpy().reshape(nd, nd, case.npz, order="F")[NG:-NG, NG:-NG]
    dl0 = d0["delp_out"].cpu().numpy().reshape(nd, nd, case.npz, order="F")[NG:-NG, NG:-NG]
    np.testing.assert_array_equal(dl1, dl0)


@pytest.mark.parametrize("form", [fa.FORM_FUSED32, fa.FORM_FUSED64, fa.FORM_PLAIN64])
def test_six_face_halos_come_from_the_neighbours(gpu_ready, form):
    dt = torch.float32 if form == fa.FORM_FUSED32 else torch.float64
    case = cases.make_case("6xC24L6Q2", n=24, npz=6, nq=2, ntiles=6, courant=1.7, delp_kind=1,
                           delp_amp=(0.05, 0.15), seed=17, with_pt=True)
    print(f"[poison] six-face q_in halos, form {form}", flush=True)
    ctx = fa.Context(case.grids, case.npz)
    d0 = _inputs(case, dt)
    _step(ctx, case, form, d0)
    d1 = _inputs(case, dt)
    nd = case.n + 2 * NG
    q = d1["q_in"].view(6, case.nq, case.npz, nd, nd)
    interior = q[..., NG:-NG, NG:-NG].clone()
    q.fill_(float("nan"))
    q[..., NG:-NG, NG:-NG] = interior
    _step(ctx, case, form, d1)
    ctx.close()
    got, want = _interior_q(case, d1["q_out"]), _interior_q(case, d0["q_out"])
    assert not np.isnan(got).any()
    np.testing.assert_array_equal(got, want)
