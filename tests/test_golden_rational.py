"""The FP64 oracle against the exact-rational restatement of hord 8 PPM
fluxes and of the kord 9 cs_profile + map1_ppm remap (tests/golden/
rational_ref.py, written from the published methods).  The committed
fixtures (tests/golden/rational_fixtures.json, made by
tests/golden/make_fixtures.py) hold seeded inputs and the float64 nearest
each exact result; their sha256 is checked before use.
"""
import json
import os
import sys
from fractions import Fraction as F

import numpy as np
import pytest

from oracle import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)

import make_fixtures  # noqa: E402
import rational_ref as rr  # noqa: E402


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(HERE, "rational_fixtures.json")) as f:
        doc = json.load(f)
    assert make_fixtures.data_hash(doc["data"]) == doc["sha256"], "fixture data changed"
    return doc["data"]


def _a(xs):
    return np.array([float.fromhex(x) for x in xs])


def test_oracle_ppm_hord8_matches_exact_rationals(fixtures):
    worst = 0.0
    for case in fixtures["ppm_hord8"]:
        q, c, d, want = _a(case["q"]), _a(case["c"]), _a(case["dxa"]), _a(case["flux"])
        for sweep in (oracle.xppm_line, oracle.yppm_line):
            got = (sweep(q, c, 8, edge=case["edge"], dxa=d) if sweep is oracle.xppm_line
                   else sweep(q, c, 8, edge=case["edge"], dya=d))
            worst = max(worst, np.abs(got - want).max() / np.abs(q).max())
    assert worst <= 1e-14, worst


def test_oracle_cs_profile_kord9_matches_exact_rationals(fixtures):
    worst = 0.0
    for case in fixtures["cs_kord9"]:
        pe1, a, pe2 = _a(case["pe1"]), _a(case["a"]), _a(case["pe2"])
        a4 = oracle.cs_profile(a[None], np.diff(pe1)[None], iv=case["iv"], kord=9)
        q2 = oracle.map1_ppm(pe1[None], a[None], pe2[None], iv=case["iv"], kord=9)[0]
        scale = np.abs(a).max()
        worst = max(worst, np.abs(a4[1, 0] - _a(case["aL"])).max() / scale,
                    np.abs(a4[2, 0] - _a(case["aR"])).max() / scale,
                    np.abs(q2 - _a(case["q2"])).max() / scale)
    assert worst <= 1e-13, worst


def test_fixtures_are_the_restatement_output(fixtures):
    # the stored results are exactly what the restatement gives on the stored
    # inputs (a few cases, to keep the suite fast)
    for case in fixtures["ppm_hord8"][::3]:
        fl = rr.ppm_flux_hord8([F(v) for v in _a(case["q"])], [F(v) for v in _a(case["c"])],
                               [F(v) for v in _a(case["dxa"])] if case["edge"] else None,
                               edge=case["edge"])
        assert [float(v).hex() for v in fl] == case["flux"]
    for case in fixtures["cs_kord9"][::4]:
        pe1 = [F(v) for v in _a(case["pe1"])]
        q2 = rr.map1_ppm_kord9(pe1, [F(v) for v in _a(case["a"])], [F(v) for v in _a(case["pe2"])],
                               case["iv"])
        assert [float(v).hex() for v in q2] == case["q2"]


def test_restatement_limiters_are_active(fixtures):
    # the fixtures exercise the limiters: some PPM cells and some layers end
    # up away from the unconstrained values
    changed = 0
    for case in fixtures["cs_kord9"]:
        pe1 = [F(v) for v in _a(case["pe1"])]
        a = [F(v) for v in _a(case["a"])]
        d = [pe1[k + 1] - pe1[k] for k in range(len(a))]
        q = rr.spline_interfaces(a, d)
        aL = [F(v) for v in _a(case["aL"])]
        changed += sum(1 for k in range(len(a)) if float(aL[k]) != float(q[k]))
    assert changed >= 10
