"""Writes tests/golden/rational_fixtures.json: seeded inputs and the exact
results of tests/golden/rational_ref.py for

* hord 8 PPM fluxes on 8-cell lines (3-cell halos), bounded lines and lines
  across a cube edge with the real gnomonic widths of a C24 tile;
* kord 9 cs_profile + map1_ppm on 6- and 10-layer columns (iv = 0 and 1).

Every number is stored as a float64 hex string (inputs exactly; results as
the float64 nearest the exact rational).  The file carries the sha256 of its
data block, which tests/test_golden_rational.py checks.

Usage: python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
from fractions import Fraction as F

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import rational_ref as rr  # noqa: E402

NG = 3
OUT = os.path.join(HERE, "rational_fixtures.json")


def hx(v):
    return float(v).hex()


def ppm_cases(rng):
    from fv3atm_amd import cases
    g = cases.make_grid(24, 1)
    out = []
    n = 8
    for t in range(12):
        edge = t >= 6
        kind = t % 3
        x = np.arange(n + 2 * NG, dtype=np.float64)
        if kind == 0:    # rough positive data
            q = rng.uniform(0.0, 1.0, n + 2 * NG)
        elif kind == 1:  # a smooth bump plus noise
            q = np.exp(-((x - rng.uniform(3, 10)) / 3.0) ** 2) + 0.05 * rng.uniform(size=x.size)
        else:            # steps
            q = np.where(rng.uniform(size=x.size) < 0.5, 1.0, 0.0) + 0.01 * x
        c = rng.uniform(-0.95, 0.95, n + 1)
        c[np.abs(c) < 0.05] = 0.3
        if edge:
            # the widths of 14 consecutive cells of a real gnomonic row,
            # rescaled so the two faces of the tile differ
            row = 3 + rng.integers(0, 24)
            d = g.dxa[:, row][: n + 2 * NG].copy()
            d = d / d.mean()
            d[:NG] *= rng.uniform(0.9, 1.1)
            d[-NG:] *= rng.uniform(0.9, 1.1)
        else:
            d = np.ones(n + 2 * NG)
        flux = rr.ppm_flux_hord8([F(v) for v in q], [F(v) for v in c],
                                 [F(v) for v in d] if edge else None, edge=edge)
        out.append(dict(q=[hx(v) for v in q], c=[hx(v) for v in c], dxa=[hx(v) for v in d],
                        edge=edge, flux=[hx(v) for v in flux]))
    return out


def column_cases(rng):
    out = []
    for t in range(40):
        iv = t % 2
        kind = (t // 2) % 5
        km, kn = (6, 5) if t < 20 else (10, 7)
        d = rng.uniform(0.5, 1.6, km)
        pe1 = np.concatenate([[0.0], np.cumsum(d)])
        pe2 = np.concatenate([[0.0], np.sort(rng.uniform(0.05, pe1[-1] - 0.05, kn - 1)), [pe1[-1]]])
        x = 0.5 * (pe1[1:] + pe1[:-1])
        if kind == 0:
            a = rng.uniform(0.0, 1.0, km)
        elif kind == 1:
            a = 1.0 + 0.4 * x + 0.05 * x * x + 0.01 * rng.uniform(size=km)
        elif kind == 2:
            a = np.where(np.arange(km) % 2 == 0, 1.0, 0.2) + 0.01 * rng.uniform(size=km)
        elif kind == 3:
            a = np.exp(-((x - rng.uniform(1, 4)) / 0.7) ** 2) + 0.02 * rng.uniform(size=km)
        else:  # steep, positive, with near-zero minima
            a = rng.uniform(0.0, 1.0, km) ** 6
        pe1f = [F(v) for v in pe1]
        # the spline is set up on the thicknesses as the oracle forms them
        q2 = rr.map1_ppm_kord9(pe1f, [F(v) for v in a], [F(v) for v in pe2], iv)
        dd = [pe1f[k + 1] - pe1f[k] for k in range(km)]
        aL, aR = rr.cs_profile_kord9([F(v) for v in a], dd, iv)
        out.append(dict(iv=iv, pe1=[hx(v) for v in pe1], a=[hx(v) for v in a],
                        pe2=[hx(v) for v in pe2], q2=[hx(v) for v in q2],
                        aL=[hx(v) for v in aL], aR=[hx(v) for v in aR]))
    return out


def data_hash(data) -> str:
    return hashlib.sha256(json.dumps(data, sort_keys=True).encode()).hexdigest()


def build():
    rng = np.random.default_rng(20240601)
    data = dict(ppm_hord8=ppm_cases(rng), cs_kord9=column_cases(rng))
    return dict(sha256=data_hash(data), data=data)


if __name__ == "__main__":
    doc = build()
    with open(OUT, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}: {len(doc['data']['ppm_hord8'])} PPM lines, "
          f"{len(doc['data']['cs_kord9'])} columns, sha256 {doc['sha256'][:16]}")
