"""GPU checks of the FP32 step's column-shared remap kernels (remap kernel 0:
a column's levels spread over the lanes of a wave; kernel 2: one thread per
(column, field)) against the one-thread-per-column FP32 remap (kernel 1),
the FP64 oracle,
and what the mathematics fixes: means of a profile linear in pressure map to
the means of the same line over the target layers, constants stay constant,
column mass is conserved.  Layer thicknesses are perturbed by up to 45 % per
cell so that Eulerian interfaces cut the Lagrangian layers at every offset,
several per layer in places.
"""
import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases
from oracle import oracle
from tests._util import interior, rel_l2

pytestmark = pytest.mark.gpu
NG = 3


def _dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(
        device="cuda", dtype=dtype)


def run32(case, remap, kord_tr, kord_tm, t_min=184.0):
    ctx = fa.Context(case.grids, case.npz)
    fa.set_launch(ctx, 0, 0, remap)
    q_in = _dev(case.q, torch.float32)
    q_out = torch.zeros_like(q_in)
    pt = _dev(case.pt, torch.float32)
    dp1 = _dev(case.dp1, torch.float64)
    delp_out = torch.zeros_like(dp1)
    cx, cy, mfx, mfy = (_dev(getattr(case, k), torch.float32) for k in ("cx", "cy", "mfx", "mfy"))
    ak, bk = _dev(case.ak, torch.float64), _dev(case.bk, torch.float64)
    fa.step_dev(ctx, fa.FORM_FUSED32, case.nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy,
                ak, bk, case.ptop, case.hord, kord_tr, kord_tm, t_min, 0)
    torch.cuda.synchronize()
    out = dict(q=q_out.double().cpu().numpy().reshape(case.q.shape, order="F"),
               pt=pt.double().cpu().numpy().reshape(case.pt.shape, order="F"),
               delp=delp_out.cpu().numpy().reshape(case.dp1.shape, order="F"))
    ctx.close()
    return out


def structured_case(courant, seed=5, n=40, npz=40, nq=5):
    case = cases.make_case(f"C{n}L{npz}Q{nq}", n=n, npz=npz, nq=nq, courant=courant,
                           delp_kind=0, delp_amp=(0.45,), seed=seed, with_pt=True)
    nd = n + 2 * NG
    rng = np.random.default_rng(seed)
    k = np.arange(npz)[None, None, :]
    base = 1.0 + 0.3 * rng.uniform(size=(nd, nd, 1))
    phase = rng.uniform(0, 2 * np.pi, size=(nd, nd, 1))
    q = np.zeros((nd, nd, npz, nq), order="F")
    q[..., 0] = base * (1.5 + np.sin(0.35 * k + phase))                       # smooth, positive
    q[..., 1] = base * ((k + (phase > np.pi)) % 9 < 4)                        # steps with zeros
    q[..., 2] = base * np.sin(0.6 * k + phase)                                # changes sign
    q[..., 3] = base * 1e-2 * np.exp(-0.25 * (npz - 1 - k))                   # 4 decades
    q[..., 4] = base * rng.uniform(0.0, 1.0, size=(nd, nd, npz)) ** 3         # noise, extrema
    case.q = np.asfortranarray(q)
    return case


def oracle_sensitivity(case, q_adv, kord_tr, kord_tm, ref):
    """Rel L2 change of the FP64 remap per tracer when its input moves by
    1e-7 relative: the floor no FP32 evaluation can beat.  kord 10 compares
    |2 a1 - (aL + aR)| with |aL - aR|, an exact tie whenever an interface value
    is clamped to the cell mean, so on rough fields that floor is ~1e-3."""
    rng = np.random.default_rng(1)
    qp = q_adv * (1 + 1e-7 * rng.standard_normal(q_adv.shape))
    q2, _, _ = oracle.remap_tracers(case, qp, case.dp1, case.pt, kord_tr=kord_tr, kord_tm=kord_tm)
    return [rel_l2(interior(q2[..., i]), interior(ref[..., i])) for i in range(case.nq)]


@pytest.mark.parametrize("kord_tr,kord_tm", [(9, 10), (10, 9), (22, 22)])
def test_fast_remap_matches_column_remap_and_oracle(gpu_ready, kord_tr, kord_tm):
    case = structured_case(courant=0.6)
    print("case", case.name, "kord", kord_tr, kord_tm, flush=True)
    fast = run32(case, 0, kord_tr, kord_tm)
    col = run32(case, 1, kord_tr, kord_tm)
    adv = oracle.tracer_2d(case, "1L")
    q64, delp64, pt64 = oracle.remap_tracers(case, adv["q"], case.dp1, case.pt, kord_tr=kord_tr,
                                             kord_tm=kord_tm)
    sens = oracle_sensitivity(case, adv["q"], kord_tr, kord_tm, q64)
    # target thickness: the same FP64 expression on both sides
    np.testing.assert_array_equal(interior(fast["delp"]), interior(col["delp"]))
    np.testing.assert_allclose(interior(fast["delp"]), interior(delp64), rtol=1e-13)
    for iq in range(case.nq):
        a, b, r = (interior(x[..., iq]) for x in (fast["q"], col["q"], q64))
        e_col, e_64 = rel_l2(a, b), rel_l2(a, r)
        e_ref = rel_l2(b, r)
        print(f"tracer {iq}: rel L2 fast-col {e_col:.2e} fast-fp64 {e_64:.2e} col-fp64 {e_ref:.2e}"
              f" fp64 sensitivity {sens[iq]:.2e}")
        tol = max(1e-5, 4.0 * sens[iq])
        if kord_tr != 10:
            assert sens[iq] <= 1e-6  # no ties: the FP32 contract applies as stated
        assert e_col <= tol and e_64 <= tol
    a, r = interior(fast["pt"]), interior(pt64)
    print(f"pt: rel L2 fast-fp64 {rel_l2(a, r):.2e}")
    assert rel_l2(a, r) <= 1e-6


# 8, 32, 63: 2 or 4 levels on each of 16 lanes (32: interface km+1 opens a
# lane); 64, 127: 4 levels on each of 32 lanes.  18 x 18 columns leave a
# partly filled last workgroup.
@pytest.mark.parametrize("npz", [8, 32, 63, 64, 127])
@pytest.mark.parametrize("kord_tr,kord_tm", [(9, 10), (10, 9), (8, 8), (22, 22)])
def test_lane_remap_layouts(gpu_ready, npz, kord_tr, kord_tm):
    case = structured_case(courant=0.6, seed=11, n=18, npz=npz)
    print("case", case.name, "kord", kord_tr, kord_tm, flush=True)
    lanes = run32(case, 0, kord_tr, kord_tm)
    per_field = run32(case, 2, kord_tr, kord_tm)
    adv = oracle.tracer_2d(case, "1L")
    q64, _, pt64 = oracle.remap_tracers(case, adv["q"], case.dp1, case.pt, kord_tr=kord_tr, kord_tm=kord_tm)
    sens = oracle_sensitivity(case, adv["q"], kord_tr, kord_tm, q64)
    np.testing.assert_array_equal(interior(lanes["delp"]), interior(per_field["delp"]))
    for iq in range(case.nq):
        a, b, r = (interior(x[..., iq]) for x in (lanes["q"], per_field["q"], q64))
        e_pf, e_64 = rel_l2(a, b), rel_l2(a, r)
        print(f"tracer {iq}: rel L2 lanes-per_field {e_pf:.2e} lanes-fp64 {e_64:.2e} "
              f"per_field-fp64 {rel_l2(b, r):.2e} sensitivity {sens[iq]:.2e}")
        tol = max(1e-5, 4.0 * sens[iq])
        assert e_pf <= tol and e_64 <= tol
    a, r = interior(lanes["pt"]), interior(pt64)
    print(f"pt: rel L2 lanes-fp64 {rel_l2(a, r):.2e}")
    assert rel_l2(a, r) <= 1e-6


def test_lane_remap_many_workgroups(gpu_ready):
    # C96 L127: 9216 columns in 576 workgroups dealt over the XCDs; the lane
    # kernel against the per-(column, field) kernel, and column mass
    case = structured_case(courant=0.6, seed=13, n=96, npz=127, nq=6)
    print("case", case.name, flush=True)
    lanes = run32(case, 0, 9, 10)
    per_field = run32(case, 2, 9, 10)
    for iq in range(case.nq):
        e = rel_l2(interior(lanes["q"][..., iq]), interior(per_field["q"][..., iq]))
        print(f"tracer {iq}: rel L2 lanes-per_field {e:.2e}")
        assert e <= 1e-6
    e = rel_l2(interior(lanes["pt"]), interior(per_field["pt"]))
    print(f"pt: rel L2 lanes-per_field {e:.2e}")
    assert e <= 1e-6


@pytest.mark.parametrize("kord", [9, 10, 22])
def test_fast_remap_linear_in_pressure_is_exact(gpu_ready, kord):
    # no transport (Courant 1e-9): the remap alone acts on q = 2 + 3 p / ps,
    # whose layer means are the line at the layer mid-pressure
    case = structured_case(courant=1e-9, seed=8)
    pe1 = case.ptop + np.concatenate([np.zeros(case.dp1.shape[:2] + (1,)),
                                      np.cumsum(case.dp1, axis=2)], axis=2)
    ps = pe1[..., -1:]
    line = lambda pe: 2.0 + 3.0 * 0.5 * (pe[..., 1:] + pe[..., :-1]) / cases.PS
    for iq in range(case.nq):
        case.q[..., iq] = line(pe1)
    got = run32(case, 0, kord, kord)
    pe2 = case.ak[None, None, :] + case.bk[None, None, :] * ps
    pe2[..., 0] = case.ptop
    pe2[..., -1] = ps[..., 0]
    want = interior(line(pe2))
    for iq in range(case.nq):
        err = np.abs(interior(got["q"][..., iq]) - want).max() / 5.0
        print(f"kord {kord} tracer {iq}: max rel error {err:.2e}")
        assert err <= 2e-6


def test_fast_remap_constancy_and_column_mass(gpu_ready):
    case = structured_case(courant=1e-9, seed=9)
    case.q[..., 0] = 0.75
    got = run32(case, 0, 9, 10)
    assert np.abs(interior(got["q"][..., 0]) - 0.75).max() <= 4e-7
    dp1, dp2 = interior(case.dp1), interior(got["delp"])
    for iq in (1, 3, 4):
        m0 = (interior(case.q[..., iq]) * dp1).sum(axis=2)
        m1 = (interior(got["q"][..., iq]) * dp2).sum(axis=2)
        rel = np.abs(m1 - m0).max() / np.abs(m0).max()
        print(f"tracer {iq}: column mass error {rel:.2e}")
        assert rel <= 2e-6
    # non-negative input stays non-negative (iv = 0 limiter)
    for iq in (1, 3, 4):
        qi = interior(got["q"][..., iq])
        assert qi.min() >= -1e-6 * qi.max()
