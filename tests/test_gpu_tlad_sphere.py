"""The tangent-linear / adjoint pair on the six-face sphere.  Every
sub-cycle starts from halos refreshed from the neighbouring faces (values and
tangents); the adjoint carries the adjoint of every halo and corner-block
read back onto the cells they were copied from, on the neighbour faces.

- the TL's value part equals the FP64 oracle's six-tile forward to 1e-12;
- dot-product test <TL dx, y> = <dx, AD y> to 1e-12 over all six faces;
- central differences of the forward match the TL (kord 0: exact to rounding);
- an adjoint seeded on one face reaches the other faces;
- the sphere workload's final step (FP64 forward + adjoint of tracer 0) at
  a reduced size: 6 x C48, 16 levels.
"""
import numpy as np
import pytest

from fv3atm_amd import cases
from oracle import oracle
from tests._util import assert_close, interior
from tests.test_gpu_tlad import NG, Runner, dot, grf

pytestmark = pytest.mark.gpu


def _sphere_smooth(case, seed):
    """Smooth positive tracers of the cell-centre position on the sphere."""
    rng = np.random.default_rng(seed)
    nd, npz, nq = case.n + 2 * NG, case.npz, case.nq
    q = np.zeros((nd, nd, npz, nq, 6), order="F")
    k = np.arange(npz)[None, None, :]
    for iq in range(nq):
        a = rng.uniform(0.5, 1.5, 3)
        p = rng.uniform(0, 2 * np.pi, 3)
        for t, g in enumerate(case.grids):
            x, y, z = (g.xyz_c[c][:, :, None] for c in range(3))
            q[..., iq, t] = 1.2 + 0.4 * np.sin(a[0] * 2 * x + p[0]) * np.cos(a[1] * 2 * y + p[1]) * \
                (1 + 0.2 * np.sin(a[2] * z + 0.2 * k + p[2]))
    return q


CASES = {
    "6xC16L8Q2-1L-jet": (True, 0, 9, lambda: cases.make_case("6xC16L8Q2", n=16, npz=8, nq=2, ntiles=6, courant=2.1,
                                                             jet=True, delp_kind=1, delp_amp=(0.05, 0.2), seed=5)),
    "6xC16L8Q2-2d": (False, 2, 10, lambda: cases.make_case("6xC16L8Q2", n=16, npz=8, nq=2, ntiles=6, courant=1.3,
                                                           delp_kind=1, delp_amp=(0.05, 0.2), seed=6, q_split=2)),
    "6xC16L6Q1-1L-noremap": (True, 0, 0, lambda: cases.make_case("6xC16L6Q1", n=16, npz=6, nq=1, ntiles=6,
                                                                 courant=1.6, delp_kind=1, delp_amp=(0.05, 0.2),
                                                                 seed=7)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_six_face_value_and_dot_product(gpu_ready, name):
    per_level, q_split, kord, mk = CASES[name]
    case = mk()
    print(name, flush=True)
    r = Runner(case, per_level, q_split, kord)
    dq = grf(case.q.shape, 1)
    q1, dq1, delp = r.tl(case.q, dq)
    adv = oracle.tracer_2d(case, "1L" if per_level else "2d")
    print("nsplit", adv["nsplit"])
    if kord:
        ref, _, _ = oracle.remap_tracers(case, adv["q"], case.dp1, None, kord_tr=kord)
    else:
        ref = adv["q"]
    assert_close(interior(q1), interior(ref), 1e-12, "TL value", floor=1e-300)
    y = grf(case.q.shape, 2)
    qb = r.ad(case.q, y)
    lhs, rhs = dot(dq1, y), dot(dq, qb)
    print(f"<TL dx, y> = {lhs:.15e}  <dx, AD y> = {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    assert np.all(qb[:NG] == 0) and np.all(qb[:, :NG] == 0) and np.all(qb[-NG:] == 0) and np.all(qb[:, -NG:] == 0)


def test_six_face_tl_matches_central_differences(gpu_ready):
    case = cases.make_case("6xC16L6Q1", n=16, npz=6, nq=1, ntiles=6, courant=1.6, delp_kind=1, delp_amp=(0.05, 0.2),
                           seed=7)
    case.q = _sphere_smooth(case, 4)
    r = Runner(case, True, 0, 0)
    dq = grf(case.q.shape, 3)
    _, dq1, _ = r.tl(case.q, dq)
    scale = np.abs(case.q).max()
    errs = {}
    for eps in (1e-6, 1e-7, 1e-8):
        qp, _, _ = r.tl(case.q + eps * scale * dq, np.zeros_like(dq))
        qm, _, _ = r.tl(case.q - eps * scale * dq, np.zeros_like(dq))
        fd = (qp - qm) / (2 * eps * scale)
        e = interior(fd - dq1)
        errs[eps] = np.linalg.norm(e) / np.linalg.norm(interior(dq1))
        per_face = [np.linalg.norm(e[..., t]) / np.linalg.norm(interior(dq1)[..., t]) for t in range(6)]
        worst = np.unravel_index(np.argmax(np.abs(e)), e.shape)
        print(f"eps {eps:g}: rel L2 |FD - TL| {errs[eps]:.2e}; per face " +
              " ".join(f"{v:.1e}" for v in per_face) + f"; worst cell {worst}")
    # a +-eps interval that crosses a monotone-limiter kink (a tie in the
    # PPM limits) moves FD off the frozen-branch TL in that cell; which faces
    # that happens on changes with eps.  At eps = 1e-8 no cell crosses one and
    # every face agrees to rounding (about 4e-8, the FD's own cancellation)
    assert errs[1e-8] <= 1e-7
    assert min(errs.values()) <= 1e-7


def test_adjoint_crosses_the_cube_edges(gpu_ready):
    # seeded on face 1 only, the adjoint reaches the faces the flow came from
    per_level, q_split, kord, mk = CASES["6xC16L6Q1-1L-noremap"]
    case = mk()
    r = Runner(case, per_level, q_split, kord)
    y = np.zeros(case.q.shape, order="F")
    y[..., 0] = grf(case.q.shape[:-1], 8)
    qb = r.ad(case.q, y)
    reach = [float(np.abs(interior(qb[..., t])).max()) for t in range(6)]
    print("max |adjoint| per face", reach)
    assert sum(v > 0 for v in reach[1:]) >= 2
    dq = grf(case.q.shape, 9)
    _, dq1, _ = r.tl(case.q, dq)
    lhs, rhs = dot(dq1, y), dot(dq, qb)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_sphere_final_step_adjoint_reduced(gpu_ready):
    # the sphere workload's final step, FP64 forward + adjoint of tracer 0:
    # solid-body rotation over the corners, kord 9 remap, 6 x C48 L16
    case = cases.make_case("6xC48L16Q1", n=48, npz=16, nq=1, ntiles=6, courant=1.8, delp_kind=1,
                           delp_amp=(0.05, 0.15), seed=13)
    print("6xC48L16Q1", flush=True)
    r = Runner(case, True, 0, 9)
    dq = grf(case.q.shape, 31)
    q1, dq1, _ = r.tl(case.q, dq)
    y = grf(case.q.shape, 32)
    qb = r.ad(case.q, y)
    lhs, rhs = dot(dq1, y), dot(dq, qb)
    print(f"<TL dx, y> = {lhs:.15e}  <dx, AD y> = {rhs:.15e}  rel {abs(lhs - rhs) / abs(lhs):.2e}")
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
