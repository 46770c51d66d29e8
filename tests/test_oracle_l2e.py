"""CPU pins of the oracle's full Lagrangian_to_Eulerian (oracle/
fv3_oracle_l2e.c): what the remap must conserve column by column (the wind
momenta on their face columns, w's momentum, the column depth through
delz), the identity when the Lagrangian and Eulerian levels coincide, the
domain's total energy under the fix-up, pkz as the layer mean of p^kappa
over log p (by quadrature), and the GFDL_interstitial kmp rule."""
import ctypes

import numpy as np
import pytest

from fv3atm_amd import _ffi, cases
from oracle import oracle

NG = 3


def _case(n=12, km=10, seed=3, ntiles=1):
    return cases.make_case(f"C{n}L{km}", n=n, npz=km, nq=2, ntiles=ntiles, delp_kind=1,
                           delp_amp=(0.05, 0.2), seed=seed, with_pt=ntiles == 1)


def _run(case, st, hydrostatic, consv=0.0, kord=9):
    areas = [g.area for g in case.grids]
    return oracle.lagrangian_to_eulerian_full(case.n, case.ntiles, case.npz, case.nq, case.ak, case.bk,
                                             case.ptop, st, kord, 10, kord, kord, hydrostatic, consv, areas,
                                             akap=cases.KAPPA, cp=cases.CP_AIR, rdgas=cases.RDGAS,
                                             grav=cases.GRAV)


def _pe(case, dp):
    nd = dp.shape[0]
    return cases.PTOP + np.concatenate([np.zeros((nd, nd, 1)), np.cumsum(dp, axis=2)], axis=2)


def _pe2(case, ps):
    p = case.ak[None, None, :] + case.bk[None, None, :] * ps[..., None]
    p[..., 0], p[..., -1] = case.ptop, ps
    return p


@pytest.mark.parametrize("kord", [9, 10])
def test_column_integrals_conserved(kord):
    case = _case()
    st = cases.l2e_state(case, 1)
    out, _ = _run(case, st, False, kord=kord)
    n, s = case.n, slice(NG, NG + case.n)
    pe1 = _pe(case, st["delp"])
    ps = pe1[..., -1]
    pe2 = _pe2(case, ps)
    # u on the y-faces of rows js..je+1: interfaces averaged over the cells j-1, j
    p1u = 0.5 * (pe1[s, NG - 1:NG + n] + pe1[s, NG:NG + n + 1])
    p2u = case.ak + 0.5 * case.bk * (ps[s, NG - 1:NG + n] + ps[s, NG:NG + n + 1])[..., None]
    m1 = (st["u"][s, NG:NG + n + 1] * np.diff(p1u, axis=2)).sum(2)
    m2 = (out["u"][s, NG:NG + n + 1] * np.diff(p2u, axis=2)).sum(2)
    np.testing.assert_allclose(m2, m1, rtol=1e-12, atol=1e-12 * np.abs(m1).max())
    p1v = 0.5 * (pe1[NG - 1:NG + n, s] + pe1[NG:NG + n + 1, s])
    p2v = case.ak + 0.5 * case.bk * (ps[NG - 1:NG + n, s] + ps[NG:NG + n + 1, s])[..., None]
    m1 = (st["v"][NG:NG + n + 1, s] * np.diff(p1v, axis=2)).sum(2)
    m2 = (out["v"][NG:NG + n + 1, s] * np.diff(p2v, axis=2)).sum(2)
    np.testing.assert_allclose(m2, m1, rtol=1e-12, atol=1e-12 * np.abs(m1).max())
    # w's column momentum and the column depth (sum of -delz)
    d1, d2 = st["delp"][s, s], np.diff(pe2[s, s], axis=2)
    np.testing.assert_allclose((out["w"][s, s] * d2).sum(2), (st["w"][s, s] * d1).sum(2), rtol=1e-11,
                               atol=1e-12 * np.abs(st["w"] * st["delp"]).sum(2).max())
    np.testing.assert_allclose(out["delz"].sum(2), st["delz"].sum(2), rtol=1e-12)
    np.testing.assert_allclose(out["delp"][s, s], d2, rtol=1e-13)


def test_identity_when_levels_coincide():
    case = _case(seed=4)
    ps = _pe(case, case.dp1)[..., -1]
    case.dp1 = np.asfortranarray(np.diff(_pe2(case, ps), axis=2))  # the Eulerian levels already
    st = cases.l2e_state(case, 2)
    out, _ = _run(case, st, False)
    s = slice(NG, NG + case.n)
    for k in ("u", "v"):
        a, b = out[k], st[k]
        sl = (s, slice(NG, NG + case.n + 1)) if k == "u" else (slice(NG, NG + case.n + 1), s)
        np.testing.assert_allclose(a[sl], b[sl], rtol=1e-13, atol=1e-13 * np.abs(b).max())
    np.testing.assert_allclose(out["w"][s, s], st["w"][s, s], rtol=1e-13, atol=1e-13 * np.abs(st["w"]).max())
    np.testing.assert_allclose(out["delz"], st["delz"], rtol=1e-13)
    np.testing.assert_allclose(out["pt"][s, s], st["pt"][s, s], rtol=1e-13)


def _total_energy(case, u, v, pt, pkz, dp):
    """sum over cells and levels of (cp pt pkz + (ua^2 + va^2)/2) dp area."""
    n, s = case.n, slice(NG, NG + case.n)
    ua = 0.5 * (u[s, NG:NG + n] + u[s, NG + 1:NG + n + 1])
    va = 0.5 * (v[NG:NG + n, s] + v[NG + 1:NG + n + 1, s])
    e = (cases.CP_AIR * pt[s, s] * pkz + 0.5 * (ua ** 2 + va ** 2)) * dp[s, s]
    return float((e.sum(2) * case.grids[0].area[s, s]).sum())


def test_energy_fixup_restores_the_total_energy():
    case = _case(seed=5)
    st = cases.l2e_state(case, 3, hydrostatic=True)
    e0 = _total_energy(case, st["u"], st["v"], st["pt"], st["pkz"], st["delp"])
    raw, _ = _run(case, st, True, consv=0.0)
    fixed, dtmp = _run(case, st, True, consv=1.0)
    e_raw = _total_energy(case, raw["u"], raw["v"], raw["pt"], raw["pkz"], raw["delp"])
    e_fix = _total_energy(case, fixed["u"], fixed["v"], fixed["pt"], fixed["pkz"], fixed["delp"])
    assert abs(e_raw - e0) > 1e-9 * abs(e0)  # the remap alone changes the energy
    assert abs(e_fix - e0) <= 1e-13 * abs(e0), (e0, e_raw, e_fix)
    assert dtmp != 0.0


def test_hydrostatic_pkz_is_the_log_p_mean_of_p_kappa():
    case = _case(seed=6)
    st = cases.l2e_state(case, 4, hydrostatic=True)
    out, _ = _run(case, st, True)
    s = slice(NG, NG + case.n)
    pe = _pe2(case, _pe(case, st["delp"])[..., -1])[s, s]
    lo, hi = np.log(pe[..., :-1]), np.log(pe[..., 1:])
    x = np.linspace(0.0, 1.0, 2001)  # Simpson on log p
    wts = np.ones_like(x)
    wts[1:-1:2], wts[2:-1:2] = 4.0, 2.0
    lp = lo[..., None] + (hi - lo)[..., None] * x
    mean = (np.exp(cases.KAPPA * lp) * wts).sum(-1) / (3.0 * (x.size - 1))
    np.testing.assert_allclose(out["pkz"], mean, rtol=1e-12)
    np.testing.assert_allclose(out["peln"].transpose(0, 2, 1), np.log(pe), rtol=1e-15)


def test_kmp_rule():
    # GFDL_interstitial: the first level whose full-level pressure
    # (ph2 - ph1) / log(ph2 / ph1) at p_ref exceeds 10 hPa
    km = 30
    ak, bk = cases.hybrid_coefs(km)
    lib = _ffi.lib()
    k = ctypes.c_int()
    assert lib.fv3_kmp(km, 1.0e5, _ffi.ptr(np.ascontiguousarray(ak)), _ffi.ptr(np.ascontiguousarray(bk)),
                       ctypes.byref(k)) == 0
    # uniform 3330 Pa layers: pfull(1) = 3230 / log(3330/100) = 921.6 hPa... the
    # top layer's log mean is 9.2 hPa < 10 hPa, the second's is 49.7 hPa
    assert k.value == 2
    assert lib.fv3_kmp(km, 1.0e3, _ffi.ptr(np.ascontiguousarray(ak)), _ffi.ptr(np.ascontiguousarray(bk)),
                       ctypes.byref(k)) == 0
    assert k.value == km  # no level reaches 10 hPa at a 10 hPa surface
