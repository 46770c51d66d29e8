"""Comparison helpers shared by the tests."""
import numpy as np

NG = 3


def max_rel(a, b, floor=0.0):
    """max over cells of |a-b| / max(|b|, floor); 0 when identical."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    assert a.shape == b.shape, (a.shape, b.shape)
    d = np.abs(a - b)
    den = np.maximum(np.abs(b), floor)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(d == 0, 0.0, d / np.where(den == 0, np.inf, den))
    r = np.where((d > 0) & (den == 0), np.inf, r)
    return float(r.max()) if r.size else 0.0


def assert_close(a, b, rtol, what, floor=0.0):
    r = max_rel(a, b, floor)
    assert r <= rtol, f"{what}: max relative difference {r:.3e} > {rtol:.1e}"


def interior(a, ng=NG):
    return a[ng:-ng, ng:-ng, ...]


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))
