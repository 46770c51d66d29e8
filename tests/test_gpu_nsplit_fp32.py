"""The FP32 form's sub-cycle counts near an integer Courant bound.  One
level's cmax is planted 2e-9 below 2 with Courant numbers that FP32 holds
exactly; the FP32 transport (through the 32-bit entry point) must sub-cycle
that level as often as the FP64 oracle does on the same numbers, because
cmax is evaluated in FP64 from the FP64 metrics in every form.  (Evaluated
in FP32, max|c| + 1 - sin_sg(5) can round up to 2 and add a sub-cycle.)"""
import ctypes

import numpy as np
import pytest

import fv3atm_amd as fa
from fv3atm_amd import _ffi, cases
from oracle import oracle

pytestmark = pytest.mark.gpu
NG = 3


def _planted_case():
    case = cases.make_case("C16L12", n=16, npz=12, nq=2, courant=0.5, delp_kind=1, delp_amp=(0.05, 0.2), seed=31)
    g = case.grids[0]
    k = 7  # a level below the top sixth
    # an interior cell (i, j) and an FP32 Courant number ax with
    # ax + 1 - sin_sg(5) below 2 in FP64 but rounding to 2 when evaluated in FP32
    for i in range(2, case.n):
        for j in range(2, case.n):
            s5 = g.sin_sg[i - 1 + NG, j - 1 + NG, 4]
            cand = [np.float32(1.0 + s5)]
            for _ in range(6):
                cand.append(np.nextafter(cand[-1], np.float32(0)))
            for ax in cand:
                c64 = float(ax) + 1.0 - s5
                c32 = float(np.float32(np.float32(ax) + np.float32(1.0)) - np.float32(s5))
                if c64 < 2.0 <= c32:
                    break
            else:
                continue
            break
        else:
            continue
        break
    cx = case.cx.copy()
    cx[:, :, k] *= 0.5 / max(np.abs(case.cx[:, :, k]).max(), np.abs(case.cy[:, :, k]).max())
    cy = case.cy.copy()
    cy[:, :, k] *= 0.5 / max(np.abs(case.cx[:, :, k]).max(), np.abs(case.cy[:, :, k]).max())
    cx[i - 1, j - 1 + NG, k] = float(ax)  # face i of cell (i, j); cx is (is:ie+1, jsd:jed)
    # every Courant number as FP32 holds it, so both forms see the same values
    case.cx = np.asfortranarray(cx.astype(np.float32).astype(np.float64))
    case.cy = np.asfortranarray(cy.astype(np.float32).astype(np.float64))
    return case, k, c64, c32


def test_fp32_nsplit_matches_fp64_at_an_integer_bound(gpu_ready):
    case, k, c64, c32 = _planted_case()
    print(f"planted level {k + 1}: cmax FP64 {c64:.12f}, evaluated in FP32 {c32:.9f}", flush=True)
    assert c64 < 2.0 <= c32  # the plant tells FP32 evaluation from FP64
    ref = oracle.tracer_2d(case, "1L")
    assert ref["nsplit"][k] == 2
    ctx = fa.Context(case.grids, case.npz)
    f32 = {n: np.asfortranarray(getattr(case, n).astype(np.float32)) for n in ("q", "dp1", "mfx", "mfy", "cx", "cy")}
    ns = np.zeros(case.npz, dtype=np.int32)
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    rc = _ffi.lib().fv3_tracer_2d_1l_r4(ctx.handle, P(f32["q"]), P(f32["dp1"]), P(f32["mfx"]), P(f32["mfy"]),
                                        P(f32["cx"]), P(f32["cy"]), case.n + 1, case.n + 1, case.npz, case.nq,
                                        case.hord, 0, case.dt, 0, 0, 0.0, 1.0,
                                        ns.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    _ffi.check(rc, "fv3_tracer_2d_1l_r4")
    ctx.close()
    np.testing.assert_array_equal(ns, ref["nsplit"])
