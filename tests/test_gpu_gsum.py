"""fv3_global_mass (Neumaier-compensated FP64 global tracer mass) against
math.fsum, the correctly rounded sum of the same FP64 products, on one tile
and six faces, for float and double tracers; on data built to cancel, where
a plain FP64 sum loses most of its digits; and run twice, bit for bit."""
import math

import numpy as np
import pytest
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases

pytestmark = pytest.mark.gpu
NG = 3
EPS = np.finfo(np.float64).eps


def _exact(q, delp, grids, nq, nt):
    """fsum of (q * delp) * area over the compute domains, per tracer."""
    out = []
    for iq in range(nq):
        terms = []
        for t in range(nt):
            qt = q[..., iq, t].astype(np.float64)[NG:-NG, NG:-NG]
            dt = delp[..., t][NG:-NG, NG:-NG]
            ar = grids[t].area[NG:-NG, NG:-NG, None]
            terms.append(((qt * dt) * ar).ravel())
        out.append((math.fsum(np.concatenate(terms)), np.concatenate(terms)))
    return out


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()


@pytest.mark.parametrize("nt,dtype", [(1, np.float64), (6, np.float32), (6, np.float64)])
def test_matches_correctly_rounded_sum(gpu_ready, nt, dtype):
    n, npz, nq = 24, 7, 3
    case = cases.make_case(f"{nt}xC{n}", n=n, npz=npz, nq=1, ntiles=nt, delp_kind=1, delp_amp=(0.05, 0.2), seed=3)
    rng = np.random.default_rng(5)
    nd = n + 2 * NG
    q = np.asfortranarray((rng.random((nd, nd, npz, nq, nt)) * np.array([1e-6, 1.0, 3e2])[None, None, None, :, None])
                          .astype(dtype))
    delp = np.asfortranarray(case.dp1.reshape(nd, nd, npz, nt, order="F"))
    ctx = fa.Context(case.grids, npz)
    got = fa.global_mass(ctx, _dev(q), nq, _dev(delp))
    again = fa.global_mass(ctx, _dev(q), nq, _dev(delp))
    ctx.close()
    np.testing.assert_array_equal(got, again)
    for iq, (ex, _) in enumerate(_exact(q, delp, case.grids, nq, nt)):
        err = abs(got[iq] - ex) / abs(ex)
        print(f"tracer {iq}: relative error {err:.2e} ({err / EPS:.2f} eps)")
        assert err <= 4 * EPS


def test_cancellation(gpu_ready):
    # large terms of both signs that nearly cancel, over a small background
    n, npz = 32, 4
    case = cases.make_case("C32L4", n=n, npz=npz, nq=1, delp_kind=1, delp_amp=(0.05, 0.2), seed=7)
    nd = n + 2 * NG
    rng = np.random.default_rng(11)
    sign = np.where(rng.random((nd, nd, npz)) < 0.5, 1.0, -1.0)
    q = np.asfortranarray((1e-3 * rng.random((nd, nd, npz)) + 1e13 * sign)[..., None, None])
    delp = np.asfortranarray(case.dp1.reshape(nd, nd, npz, 1, order="F"))
    ctx = fa.Context(case.grids, npz)
    got = fa.global_mass(ctx, _dev(q), 1, _dev(delp))[0]
    ctx.close()
    ex, terms = _exact(q, delp, case.grids, 1, 1)[0]
    plain = float(np.sum(terms))
    scale = np.abs(terms).sum()
    err, err_plain = abs(got - ex), abs(plain - ex)
    print(f"exact {ex:.6e}, compensated error {err:.2e}, plain sum error {err_plain:.2e}, sum|p| {scale:.2e}")
    assert err <= 4 * EPS * abs(ex) + 4 * EPS * EPS * scale * terms.size
    assert err_plain > 100 * err  # the test tells the two apart
