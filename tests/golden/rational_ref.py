"""Exact-rational (fractions.Fraction) restatement of two pieces of the path,
written from the published methods, independent of the C oracle's code:

* hord 8 PPM face flux along one grid line (xppm/yppm): Lin's monotone slope
  (Lin 2004, MWR 132, eq. 2-3), the 4th-order face value built from it, the
  perturbations limited to |2 dm| (Lin 2004's hord 8 constraint), and the
  upwind flux as the exact mean of the cell parabola over the swept part of
  the cell (Colella & Woodward 1984).  At a cube edge (Putman & Lin 2007) the
  edge face value is the mean of two one-sided, width-weighted linear
  extrapolations clipped to the four neighbouring means; the faces next to it
  take the one-sided 3-point values (11 q_far + 3 q_near)/14 +- 4/7 dm, and
  the three cells on each side pass the Colella-Woodward monotonicity
  constraint.
* kord 9 cs_profile + map1_ppm on one column (Harris et al. 2021, FV3
  scientific documentation, sec. 7): interface values from the C1
  volume-conserving parabolic spline on the non-uniform layers, closed at the
  top and bottom by the linear-exact one-sided relations of the two end
  layers; the large-scale interface bounds; extremum flags; the end-layer
  monotone limiters; for interior layers the kord 9 rule (flatten a flagged
  2-delta-p wave, else Huynh's 2nd constraint when the parabola is not
  monotone) and, for iv = 0, the positivity limiter; then each target layer
  is the exact integral of the piecewise parabola over it.

The system of interface values is solved by exact Gaussian elimination on the
full matrix, and the flux and the layer integrals come from polynomial
antiderivatives, so none of the oracle's recurrences or closed forms is
reused.  TEST INFRASTRUCTURE ONLY: make_fixtures.py uses it to write the
committed fixtures, and tests/test_golden_rational.py re-runs it.
"""
from __future__ import annotations

from fractions import Fraction as F

NG = 3


def _sgn(x):
    """Fortran sign(1, x): +1 for x >= 0."""
    return 1 if x >= 0 else -1


# --------------------------------------------------------------- PPM (hord 8)
def _parabola_mean(aL, aR, a6, s0, s1):
    """Mean over [s0, s1] of q(s) = aL + s (aR - aL + a6 (1 - s)); for an
    empty interval the value at s0."""
    b1 = aR - aL + a6

    def P(s):  # antiderivative
        return aL * s + b1 * s * s / 2 - a6 * s * s * s / 3

    if s1 == s0:
        return aL + s0 * (aR - aL + a6 * (1 - s0))
    return (P(s1) - P(s0)) / (s1 - s0)


def _cw_monotone(a, aL, aR):
    """Colella & Woodward (1984) eq. 1.10 monotonicity constraint."""
    if (aR - a) * (a - aL) <= 0:
        return a, a
    da, a6 = aR - aL, 6 * a - 3 * (aL + aR)
    if da * a6 > da * da:
        aL = 3 * a - 2 * aR
    elif da * a6 < -(da * da):
        aR = 3 * a - 2 * aL
    return aL, aR


def ppm_flux_hord8(q, c, dxa=None, edge=False):
    """q: cell means of cells 1-NG .. n+NG (list of Fractions); c: Courant
    numbers at faces 1 .. n+1 (face i between cells i-1 and i); dxa: cell
    widths (only for edge=True).  Returns the face fluxes (values) at faces
    1 .. n+1."""
    n = len(q) - 2 * NG
    Q = {i: q[i - (1 - NG)] for i in range(1 - NG, n + NG + 1)}
    W = None if dxa is None else {i: dxa[i - (1 - NG)] for i in range(1 - NG, n + NG + 1)}
    dm = {}
    for i in range(-1, n + 3):
        nb = (Q[i - 1], Q[i], Q[i + 1])
        slope = (Q[i + 1] - Q[i - 1]) / 4
        dm[i] = _sgn(slope) * min(abs(slope), max(nb) - Q[i], Q[i] - min(nb))
    # face values: centred 4th-order form from the limited slopes
    al = {i: (Q[i - 1] + Q[i]) / 2 + (dm[i - 1] - dm[i]) / 3 for i in range(0, n + 3)}
    lo, hi = (3, n - 2) if edge else (0, n + 1)  # cells that use the centred form
    aL, aR = {}, {}
    for i in range(lo, hi + 1):
        t = 2 * dm[i]
        aL[i] = Q[i] - _sgn(t) * min(abs(t), abs(al[i] - Q[i]))
        aR[i] = Q[i] + _sgn(t) * min(abs(t), abs(al[i + 1] - Q[i]))
    if edge:
        def edge_face(e):  # face e between cells e-1 (one face) and e (the other)
            left = Q[e - 1] + (Q[e - 1] - Q[e - 2]) * W[e - 1] / (W[e - 2] + W[e - 1])
            right = Q[e] - (Q[e + 1] - Q[e]) * W[e] / (W[e] + W[e + 1])
            nb = (Q[e - 2], Q[e - 1], Q[e], Q[e + 1])
            return min(max((left + right) / 2, min(nb)), max(nb))

        # west edge: faces 0, 1 (edge), 2; cells 0, 1, 2 limited
        f0 = (11 * Q[-1] + 3 * Q[0]) / 14 + F(4, 7) * dm[-1]
        f1 = edge_face(1)
        f2 = (3 * Q[1] + 11 * Q[2]) / 14 - F(4, 7) * dm[2]
        for i, (l, r) in zip((0, 1, 2), ((f0, f1), (f1, f2), (f2, al[3]))):
            aL[i], aR[i] = _cw_monotone(Q[i], l, r)
        # east edge: the mirror image; the edge face is n+1 (between cells n
        # and n+1), the 3-point faces n and n+2, cells n-1, n, n+1 limited
        g1 = (11 * Q[n - 1] + 3 * Q[n]) / 14 + F(4, 7) * dm[n - 1]
        g2 = edge_face(n + 1)
        g3 = (3 * Q[n + 1] + 11 * Q[n + 2]) / 14 - F(4, 7) * dm[n + 2]
        for i, (l, r) in zip((n - 1, n, n + 1), ((al[n - 1], g1), (g1, g2), (g2, g3))):
            aL[i], aR[i] = _cw_monotone(Q[i], l, r)
    out = []
    for i in range(1, n + 2):
        ci = c[i - 1]
        if ci > 0:  # swept part [1 - c, 1] of cell i-1
            j = i - 1
            out.append(_parabola_mean(aL[j], aR[j], 6 * Q[j] - 3 * (aL[j] + aR[j]), 1 - ci, F(1)))
        else:       # swept part [0, -c] of cell i
            out.append(_parabola_mean(aL[i], aR[i], 6 * Q[i] - 3 * (aL[i] + aR[i]), F(0), -ci))
    return out


# ------------------------------------------------------- cs_profile (kord 9)
def _solve(A, b):
    """Exact Gaussian elimination with partial pivoting (Fractions)."""
    m = len(b)
    A = [row[:] + [b[r]] for r, row in enumerate(A)]
    for col in range(m):
        piv = max(range(col, m), key=lambda r: abs(A[r][col]))
        A[col], A[piv] = A[piv], A[col]
        for r in range(col + 1, m):
            f = A[r][col] / A[col][col]
            if f:
                A[r] = [x - f * y for x, y in zip(A[r], A[col])]
    x = [F(0)] * m
    for r in range(m - 1, -1, -1):
        x[r] = (A[r][m] - sum(A[r][k] * x[k] for k in range(r + 1, m))) / A[r][r]
    return x


def spline_interfaces(a, d):
    """Interface values q_1 .. q_{km+1} of the C1 volume-conserving parabolic
    spline for layer means a and thicknesses d (top-down)."""
    km = len(a)
    A = [[F(0)] * (km + 1) for _ in range(km + 1)]
    b = [F(0)] * (km + 1)
    # C1 at the interface between layers k-1 and k (0-based k = 1..km-1):
    # (4 q_k + 2 q_{k-1} - 6 a_{k-1}) / d_{k-1} = (6 a_k - 4 q_k - 2 q_{k+1}) / d_k
    for k in range(1, km):
        A[k][k - 1] = 2 / d[k - 1]
        A[k][k] = 4 / d[k - 1] + 4 / d[k]
        A[k][k + 1] = 2 / d[k]
        b[k] = 6 * a[k - 1] / d[k - 1] + 6 * a[k] / d[k]
    # end closures (linear-exact): with g the ratio of the second layer's
    # thickness to the end layer's,
    # g (g + 1/2) q_end + (1 + g (g + 3/2)) q_next = 2 g (g + 1) a_end + a_second
    g = d[1] / d[0]
    A[0][0], A[0][1] = g * (g + F(1, 2)), 1 + g * (g + F(3, 2))
    b[0] = 2 * g * (g + 1) * a[0] + a[1]
    g = d[km - 2] / d[km - 1]
    A[km][km], A[km][km - 1] = g * (g + F(1, 2)), 1 + g * (g + F(3, 2))
    b[km] = 2 * g * (g + 1) * a[km - 1] + a[km - 2]
    return _solve(A, b)


def _end_limit(a, aL, aR, flag_only):
    """The end layers' limiters: flatten an extremum (``flag_only``: the
    extremum flag decides), else the Colella-Woodward overshoot fix."""
    if flag_only is not None:
        if flag_only:
            return a, a
    elif (a - aL) * (a - aR) >= 0:
        return a, a
    da, a6 = aR - aL, 3 * (2 * a - (aL + aR))
    if a6 * da < -(da * da):
        a6 = 3 * (aL - a)
        aR = aL - a6
    elif a6 * da > da * da:
        a6 = 3 * (aR - a)
        aL = aR - a6
    return aL, aR


def _positive(a, aL, aR):
    """iv = 0 positivity limiter: a parabola whose minimum lies below zero
    inside the layer is flattened or made one-sided."""
    a6 = 3 * (2 * a - (aL + aR))
    if a <= 0:
        return a, a
    if abs(aR - aL) < -a6 and a + (aR - aL) ** 2 / (4 * a6) + a6 / 12 < 0:
        if a < aR and a < aL:
            return a, a
        if aR > aL:
            a6 = 3 * (aL - a)
            return aL, aL - a6
        a6 = 3 * (aR - a)
        return aR - a6, aR
    return aL, aR


def cs_profile_kord9(a, d, iv):
    """Edge values (aL, aR) of every layer after the kord 9 constraints;
    iv = 0 (tracers, positive) or 1 (scalars)."""
    km = len(a)
    q = spline_interfaces(a, d)
    gam = [None] + [a[k] - a[k - 1] for k in range(1, km)]  # gam[k] = a_k - a_{k-1}
    lo_hi = lambda x, y: (min(x, y), max(x, y))
    # large-scale bounds: interfaces next to the end layers are clipped to the
    # adjacent means; inner ones too unless a neighbour mean is an extremum,
    # then only the side away from the extremum is bounded
    lo, hi = lo_hi(a[0], a[1])
    q[1] = min(max(q[1], lo), hi)
    for k in range(2, km - 1):  # interface between layers k-1 and k (0-based)
        lo, hi = lo_hi(a[k - 1], a[k])
        if gam[k - 1] * gam[k + 1] > 0:
            q[k] = min(max(q[k], lo), hi)
        elif gam[k - 1] > 0:  # a local maximum among the four means
            q[k] = max(q[k], lo)
        else:                 # a local minimum
            q[k] = min(q[k], hi)
            if iv == 0:
                q[k] = max(F(0), q[k])
    lo, hi = lo_hi(a[km - 2], a[km - 1])
    q[km - 1] = min(max(q[km - 1], lo), hi)
    aL, aR = q[:km], q[1:]
    ext = []
    for k in range(km):
        if k in (0, km - 1):
            ext.append((aL[k] - a[k]) * (aR[k] - a[k]) > 0)
        else:
            ext.append(gam[k] * gam[k + 1] < 0)
    aL, aR = list(aL), list(aR)
    # top layer (its outer value non-negative for tracers), then layer 2
    if iv == 0:
        aL[0] = max(F(0), aL[0])
    aL[0], aR[0] = _end_limit(a[0], aL[0], aR[0], None)
    aL[1], aR[1] = _end_limit(a[1], aL[1], aR[1], ext[1])
    # interior layers
    for k in range(2, km - 2):
        if ext[k] and (ext[k - 1] or ext[k + 1]):
            aL[k] = aR[k] = a[k]
        else:
            a6 = 6 * a[k] - 3 * (aL[k] + aR[k])
            if abs(a6) > abs(aL[k] - aR[k]):  # not monotone: Huynh's 2nd constraint
                p1 = a[k] - 2 * gam[k + 1]
                l1 = p1 + F(3, 2) * gam[k + 2]
                aL[k] = min(max(aL[k], min(a[k], p1, l1)), max(a[k], p1, l1))
                p2 = a[k] + 2 * gam[k]
                l2 = p2 - F(3, 2) * gam[k - 1]
                aR[k] = min(max(aR[k], min(a[k], p2, l2)), max(a[k], p2, l2))
        if iv == 0:
            aL[k], aR[k] = _positive(a[k], aL[k], aR[k])
    # bottom two layers
    if iv == 0:
        aR[km - 1] = max(F(0), aR[km - 1])
    aL[km - 2], aR[km - 2] = _end_limit(a[km - 2], aL[km - 2], aR[km - 2], ext[km - 2])
    aL[km - 1], aR[km - 1] = _end_limit(a[km - 1], aL[km - 1], aR[km - 1], None)
    return aL, aR


def map1_ppm_kord9(pe1, a, pe2, iv):
    """Target layer means for source interfaces pe1, means a, target
    interfaces pe2 (same top and bottom): exact integrals of the kord 9
    piecewise parabola."""
    km = len(a)
    d = [pe1[k + 1] - pe1[k] for k in range(km)]
    aL, aR = cs_profile_kord9(a, d, iv)

    def mass(k, pa, pb):  # integral over [pa, pb] inside layer k
        s0, s1 = (pa - pe1[k]) / d[k], (pb - pe1[k]) / d[k]
        a6 = 6 * a[k] - 3 * (aL[k] + aR[k])
        if s1 == s0:
            return F(0)
        return _parabola_mean(aL[k], aR[k], a6, s0, s1) * (pb - pa)

    out = []
    for t in range(len(pe2) - 1):
        pa, pb = pe2[t], pe2[t + 1]
        tot = F(0)
        for k in range(km):
            lo, hi = max(pa, pe1[k]), min(pb, pe1[k + 1])
            if hi > lo:
                tot += mass(k, lo, hi)
        out.append(tot / (pb - pa))
    return out
