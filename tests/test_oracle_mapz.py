"""CPU pins of the remap half of the FP64 oracle (fv_mapz restatement):
C1 continuity of the unlimited spline, identity and refine/coarsen
round trips, column conservation, constancy, monotonicity, positivity,
fillz conservation and the Lagrangian_to_Eulerian pressure bookkeeping.
"""
import numpy as np
import pytest

from fv3atm_amd import cases
from oracle import oracle

NG = 3


def _column(ncol=7, km=24, seed=0):
    rng = np.random.default_rng(seed)
    dp = rng.uniform(0.5, 1.5, (ncol, km))
    pe = np.concatenate([np.zeros((ncol, 1)), np.cumsum(dp, 1)], 1)
    return rng, dp, pe


def test_unlimited_spline_is_c1():
    # kord > 16 leaves the interface solve unconstrained: the piecewise parabola
    # (left, right, mean) must have a continuous first derivative at every
    # interior interface -- the volume-conserving spline condition.
    rng, dp, pe = _column()
    a1 = rng.uniform(0, 1, dp.shape)
    a4 = oracle.cs_profile(a1, dp, iv=1, kord=22)
    qL, qR, a = a4[1], a4[2], a1
    d_right_end = (4 * qR + 2 * qL - 6 * a) / dp          # f'(top of layer -> bottom edge)
    d_left_end = (6 * a - 4 * qL - 2 * qR) / dp
    mismatch = d_right_end[:, :-1] - d_left_end[:, 1:]
    assert np.abs(mismatch).max() <= 1e-11 * np.abs(d_left_end).max()
    # interfaces are single valued
    np.testing.assert_array_equal(qR[:, :-1], qL[:, 1:])


def test_spline_exact_for_linear_profiles():
    rng, dp, pe = _column(seed=1)
    a1 = 0.5 * (pe[:, 1:] + pe[:, :-1]) * 0.3 + 2.0  # means of 0.3 p + 2
    for kord in (9, 10, 22):
        a4 = oracle.cs_profile(a1, dp, iv=1, kord=kord)
        np.testing.assert_allclose(a4[1], 0.3 * pe[:, :-1] + 2.0, rtol=1e-13)
        np.testing.assert_allclose(a4[2], 0.3 * pe[:, 1:] + 2.0, rtol=1e-13)


@pytest.mark.parametrize("kord", [9, 10, 22])
def test_identity_remap(kord):
    rng, dp, pe = _column(seed=2)
    q1 = rng.uniform(0, 1, dp.shape)
    q2 = oracle.map1_ppm(pe, q1, pe, iv=0, kord=kord)
    np.testing.assert_allclose(q2, q1, rtol=0, atol=4e-16)


@pytest.mark.parametrize("kord,iv", [(9, 0), (10, 1), (9, 1)])
def test_conservation_constancy_refine(kord, iv):
    rng, dp, pe = _column(seed=3)
    ncol, km = dp.shape
    q1 = rng.uniform(0, 1, dp.shape)
    pe2 = np.zeros((ncol, 18))
    for i in range(ncol):
        pe2[i] = np.concatenate([[0], np.sort(rng.uniform(0, pe[i, -1], 16)), [pe[i, -1]]])
    q2 = oracle.map1_ppm(pe, q1, pe2, iv=iv, kord=kord)
    m1 = (q1 * dp).sum(1)
    m2 = (q2 * np.diff(pe2, axis=1)).sum(1)
    assert np.abs(m2 - m1).max() <= 1e-14 * m1.max()
    c2 = oracle.map1_ppm(pe, np.full(dp.shape, 3.7), pe2, iv=iv, kord=kord)
    assert np.abs(c2 - 3.7).max() <= 1e-14
    # split every layer in two and merge back: the means come back
    pef = np.zeros((ncol, 2 * km + 1))
    pef[:, 0::2] = pe
    pef[:, 1::2] = 0.5 * (pe[:, :-1] + pe[:, 1:])
    qf = oracle.map1_ppm(pe, q1, pef, iv=iv, kord=kord)
    dpf = np.diff(pef, axis=1)
    back = (qf[:, 0::2] * dpf[:, 0::2] + qf[:, 1::2] * dpf[:, 1::2]) / dp
    assert np.abs(back - q1).max() <= 1e-14


def test_positive_definite_iv0():
    # iv = 0 (tracers) keeps non-negative data non-negative; iv = 1 may
    # extrapolate below zero in the top layer, so only iv = 0 is checked
    rng, dp, pe = _column(seed=4)
    ncol, km = dp.shape
    pe2 = np.zeros((ncol, 20))
    for i in range(ncol):
        pe2[i] = np.concatenate([[0], np.sort(rng.uniform(0, pe[i, -1], 18)), [pe[i, -1]]])
    qm = np.sort(rng.uniform(0, 1, dp.shape), axis=1)
    assert oracle.map1_ppm(pe, qm, pe2, iv=0, kord=9).min() >= 0.0
    qp = rng.uniform(0, 1, dp.shape) ** 6  # steep, positive
    q3 = oracle.map1_ppm(pe, qp, pe2, iv=0, kord=9)
    assert q3.min() >= 0.0


def test_fillz_conserves_and_removes_negatives():
    rng = np.random.default_rng(5)
    im, km, nq = 9, 16, 3
    dp = rng.uniform(0.5, 1.5, (im, km))
    q = rng.uniform(0, 1, (im, km, nq))
    q[rng.uniform(size=q.shape) < 0.15] *= -0.3
    out = oracle.fillz(np.asfortranarray(q), dp)
    m0 = (q * dp[:, :, None]).sum(1)
    m1 = (out * dp[:, :, None]).sum(1)
    np.testing.assert_allclose(m1, m0, rtol=1e-13)
    assert out.min() >= 0.0


def test_lagrangian_to_eulerian_pressure_bookkeeping():
    case = cases.make_case("C10L12Q7", n=10, npz=12, nq=7, delp_kind=1, delp_amp=(0.05, 0.2),
                           seed=6, with_pt=True)
    q, delp, pt = oracle.remap_tracers(case, case.q, case.dp1, case.pt)
    # target thickness is ak/bk at the source surface pressure
    ps = cases.PTOP + case.dp1[NG:-NG, NG:-NG].sum(axis=2)
    pe2 = case.ak[None, None, :] + case.bk[None, None, :] * ps[..., None]
    pe2[..., 0] = case.ptop
    pe2[..., -1] = ps
    np.testing.assert_allclose(delp[NG:-NG, NG:-NG], np.diff(pe2, axis=2), rtol=1e-12)
    # column tracer mass conserved (mapn_tracer branch, nq > 5)
    m0 = (case.q[NG:-NG, NG:-NG] * case.dp1[NG:-NG, NG:-NG, :, None]).sum(axis=2)
    m1 = (q[NG:-NG, NG:-NG] * delp[NG:-NG, NG:-NG, :, None]).sum(axis=2)
    np.testing.assert_allclose(m1, m0, rtol=1e-13)
    # halo columns untouched
    np.testing.assert_array_equal(q[:NG], case.q[:NG])
