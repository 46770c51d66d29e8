"""torch.autograd face of the FP64 advection + tracer remap step, for the
validation harness only: the forward is the step (the value part of
fv3_step_tl_dev, the FP64 forward with every limiter branch, upwind choice
and source-layer search taken at the trajectory), the backward is
fv3_step_ad_dev, which recomputes every sub-cycle from checkpointed states.

Only the interior tracer values are differentiable: the single tile's halo
is frozen, as in the C-ABI, and its cells carry no gradient.
"""
import torch

import fv3atm_amd as fa

NG = 3


class StepSetup:
    """Fixed inputs of the step on one context: winds, thicknesses, pressure
    coefficients, the frozen halo, and the sub-cycling and remap choices."""

    def __init__(self, ctx, case, per_level, q_split, kord_tr, q_halo):
        dev = lambda a: torch.from_numpy(a.ravel(order="F").copy()).cuda()
        self.ctx, self.case = ctx, case
        self.per_level, self.q_split, self.kord = per_level, q_split, kord_tr
        self.w = {k: dev(getattr(case, k)) for k in ("dp1", "cx", "cy", "mfx", "mfy", "ak", "bk")}
        self.nid = case.n + 2 * NG
        self.q_halo = dev(q_halo)  # the full (halo included) state the interior is written into

    def view(self, flat):
        c = self.case
        return flat.view(c.nq, c.npz, self.nid, self.nid)

    def full(self, q_int):
        q = self.q_halo.clone()
        self.view(q)[:, :, NG:-NG, NG:-NG] = q_int
        return q


class Step(torch.autograd.Function):
    """q1_interior = step(q0_interior); backward by the adjoint kernels."""

    @staticmethod
    def forward(fctx, q_int, setup):
        c, w = setup.case, setup.w
        q0 = setup.full(q_int.detach())
        q1, dq1 = torch.zeros_like(q0), torch.zeros_like(q0)
        fa.step_tl_dev(setup.ctx, setup.per_level, setup.q_split, c.nq, q0, torch.zeros_like(q0), q1, dq1,
                       w["dp1"], None, w["cx"], w["cy"], w["mfx"], w["mfy"], w["ak"], w["bk"], c.ptop,
                       c.hord, setup.kord)
        fctx.setup = setup
        fctx.save_for_backward(q0)
        return setup.view(q1)[:, :, NG:-NG, NG:-NG].clone()

    @staticmethod
    def backward(fctx, g):
        setup = fctx.setup
        c, w = setup.case, setup.w
        (q0,) = fctx.saved_tensors
        qbar1 = torch.zeros_like(q0)
        setup.view(qbar1)[:, :, NG:-NG, NG:-NG] = g
        qbar0 = torch.zeros_like(q0)
        fa.step_ad_dev(setup.ctx, setup.per_level, setup.q_split, c.nq, q0, qbar1, qbar0, w["dp1"],
                       w["cx"], w["cy"], w["mfx"], w["mfy"], w["ak"], w["bk"], c.ptop, c.hord, setup.kord)
        return setup.view(qbar0)[:, :, NG:-NG, NG:-NG].clone(), None
