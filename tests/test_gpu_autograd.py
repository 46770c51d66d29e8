"""The tangent-linear / adjoint pair behind a torch.autograd.Function
(tests/_autograd_step.py, the validation harness): torch.autograd.gradcheck
of the adjoint against finite differences of the forward on smooth fields,
and the gradient of J = sum(q^2 area delp) through autograd against the
same directional derivative taken with the tangent-linear."""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases
from tests._autograd_step import NG, Step, StepSetup
from tests.test_gpu_tlad import smooth_q

pytestmark = pytest.mark.gpu


def _setup(case, per_level, q_split, kord):
    ctx = fa.Context(case.grids, case.npz)
    return StepSetup(ctx, case, per_level, q_split, kord, case.q)


def _interior(setup, q):
    return setup.view(torch.from_numpy(q.ravel(order="F").copy()).cuda())[:, :, NG:-NG, NG:-NG].clone()


@pytest.mark.parametrize("per_level,q_split,kord", [(False, 2, 0), (True, 0, 9)])
def test_gradcheck(gpu_ready, per_level, q_split, kord):
    case = cases.make_case("C8L6Q2", n=8, npz=6, nq=2, courant=0.6, delp_kind=1, delp_amp=(0.05, 0.15),
                           seed=13, q_split=q_split)
    case.q = smooth_q(case, 6)
    if kord:  # no vertical extrema: the remap's interface bounds and limiters stay inactive
        k = np.arange(case.npz)[None, None, :, None]
        case.q = np.asfortranarray(1.5 + (case.q[:, :, :1, :] - 1.2) * (1.0 + 0.15 * k))
    s = _setup(case, per_level, q_split, kord)
    q = _interior(s, case.q).requires_grad_(True)
    print(f"gradcheck per_level={per_level} q_split={q_split} kord={kord} ({q.numel()} inputs)", flush=True)
    assert torch.autograd.gradcheck(lambda x: Step.apply(x, s), (q,), eps=1e-6, atol=1e-7, rtol=1e-5,
                                    fast_mode=False)
    s.ctx.close()


def test_cost_gradient_through_autograd(gpu_ready):
    case = cases.make_case("C24L12Q3", n=24, npz=12, nq=3, courant=0.9, delp_kind=1, delp_amp=(0.05, 0.2),
                           seed=5, q_split=3)
    case.q = smooth_q(case, 7)
    s = _setup(case, False, 3, 9)
    q = _interior(s, case.q).requires_grad_(True)
    q1 = Step.apply(q, s)
    ak, bk = case.ak, case.bk
    ps = case.ptop + case.dp1.sum(axis=2)
    pe2 = ak[None, None, :] + bk[None, None, :] * ps[..., None]
    pe2[..., 0], pe2[..., -1] = case.ptop, ps
    delp = np.diff(pe2, axis=2)[NG:-NG, NG:-NG]  # (i, j, k) target thickness
    w = torch.from_numpy(np.ascontiguousarray(
        (case.grids[0].area[NG:-NG, NG:-NG, None] * delp).transpose(2, 1, 0))).cuda()
    J = (q1 ** 2 * w[None]).sum()
    J.backward()
    # the directional derivative along v by the tangent-linear
    rng = np.random.default_rng(3)
    v = torch.from_numpy(rng.standard_normal(q.shape)).cuda()
    c = case
    q0 = s.full(q.detach())
    dq0 = torch.zeros_like(q0)
    s.view(dq0)[:, :, NG:-NG, NG:-NG] = v
    q1f, dq1f = torch.zeros_like(q0), torch.zeros_like(q0)
    fa.step_tl_dev(s.ctx, False, 3, c.nq, q0, dq0, q1f, dq1f, s.w["dp1"], None, s.w["cx"], s.w["cy"],
                   s.w["mfx"], s.w["mfy"], s.w["ak"], s.w["bk"], c.ptop, c.hord, 9)
    dq1 = s.view(dq1f)[:, :, NG:-NG, NG:-NG]
    tl = float((2 * q1.detach() * w[None] * dq1).sum())
    ad = float((q.grad * v).sum())
    print(f"dJ/dv: autograd {ad:.15e} tangent-linear {tl:.15e} rel {abs(ad - tl) / abs(tl):.2e}")
    assert abs(ad - tl) <= 1e-12 * abs(tl)
    s.ctx.close()
