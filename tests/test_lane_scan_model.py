"""CPU model of the lane remap kernel's interface solve (mapz_fast.hip,
k_remap_rows): the Thomas recurrences of cs_profile/scalar_profile run as
per-lane recurrences composed by affine scans over the lanes of a column,
with the kernel's cross-lane moves restated lane by lane (DPP row_shr /
row_shl within 16-lane rows, row_bcast:15 into odd rows, wave_shr / wave_shl
across the wave, bound_ctrl reading 0 or keeping the old value, and the one
ds_bpermute of the 32-lane backward scan).  For every layout the kernel uses
(M levels on each of L lanes, two or four columns per wave) and column heights
that put interface km+1 at every slot, the scanned result must equal the
sequential recurrence: a wrong scan step, slope or carry fails it.  This pins
the scan design; the GPU tests pin the kernel's results against the oracle."""
import numpy as np
import pytest


def dpp(v, ctrl, old, bound_zero, row_mask=0xF):
    out = np.empty_like(v)
    for lane in range(64):
        row = lane // 16
        if not (row_mask >> row) & 1:
            out[lane] = old
            continue
        if 0x111 <= ctrl <= 0x11F:  # row_shr:n
            src = lane - (ctrl - 0x110)
            ok = src >= 0 and src // 16 == row
        elif 0x101 <= ctrl <= 0x10F:  # row_shl:n
            src = lane + (ctrl - 0x100)
            ok = src // 16 == row
        elif ctrl == 0x138:  # wave_shr:1
            src, ok = lane - 1, lane >= 1
        elif ctrl == 0x130:  # wave_shl:1
            src, ok = lane + 1, lane < 63
        elif ctrl == 0x142:  # row_bcast:15
            src, ok = (row - 1) * 16 + 15, row >= 1
        else:
            raise ValueError(ctrl)
        out[lane] = v[src] if ok else (0.0 if bound_zero else old)
    return out


def lane_solve(M, L, km, a, d4, rbt, gg, bot):
    """The kernel's solve for 64/L columns: returns the interface values
    Q(k), k = 1..L*M, per (column, level)."""
    lane = np.arange(64)
    g = lane % L
    nb_lo = (lambda x: dpp(x, 0x111, 0, True)) if L == 16 else (lambda x: dpp(x, 0x138, 0, True))
    nb_hi = (lambda x: dpp(x, 0x101, 0, True)) if L == 16 else (lambda x: dpp(x, 0x130, 0, True))
    at = lambda arr, i: np.array([arr[l // L, (l % L) * M + i] for l in range(64)])
    q = [at(a, i) for i in range(M)]
    D4 = [at(d4, i) for i in range(M)]
    R = [at(rbt, i) for i in range(M)]
    G = [at(gg, i) for i in range(M)]

    def slopes(sl, lo):
        s = []
        for n in (1, 2, 4, 8):
            s.append(sl.copy())
            sl = sl * dpp(sl, (0x110 if lo else 0x100) + n, 1.0, False)
        s.append(sl.copy())
        return s

    sf, sb = np.ones(64), np.ones(64)
    for i in range(M):
        sf, sb = sf * -R[i], sb * -G[i]
    fs, bs = slopes(sf, True), slopes(sb, False)

    def scan_lo(x):
        for j, n in enumerate((1, 2, 4, 8)):
            x = fs[j] * dpp(x, 0x110 + n, 0, True) + x
        if L == 32:
            x = fs[4] * dpp(x, 0x142, 0.0, True, 0xA) + x
        return x

    def scan_hi(x):
        for j, n in enumerate((1, 2, 4, 8)):
            x = bs[j] * dpp(x, 0x100 + n, 0, True) + x
        if L == 32:
            u = np.array([x[(l & 32) | 16] for l in range(64)])
            x = np.where((lane & 16) == 0, bs[4] * u + x, x)
        return x

    am1 = nb_lo(q[M - 1])
    t1 = D4[0] * q[0] + q[1]
    e = np.where(g == 0, t1, 3 * (D4[0] * q[0] + am1)) * R[0]
    for i in range(1, M):
        e = (3 * (D4[i] * q[i] + q[i - 1]) - e) * R[i]
    cin = nb_lo(scan_lo(e))
    e = np.where(g == 0, t1 * R[0], (3 * (D4[0] * q[0] + am1) - cin) * R[0])
    x = [e]
    for i in range(1, M):
        e = (3 * (D4[i] * q[i] + q[i - 1]) - e) * R[i]
        x.append(e)
    s_km, s_kb, g_kb = (km - 1) % M, km % M, km // M
    a_km1 = q[s_km - 1] if s_km > 0 else am1
    B = [np.array([bot[l // L, m] for l in range(64)]) for m in range(3)]
    qb = (B[0] * q[s_km] + a_km1 - B[1] * x[s_km]) * B[2]
    if s_kb == 0:
        qb = nb_lo(qb)
    x[s_kb] = np.where(g == g_kb, qb, x[s_kb])
    e = np.zeros(64)
    for i in range(M - 1, -1, -1):
        e = -G[i] * e + x[i]
    e = nb_hi(scan_hi(e))
    out = np.zeros((64 // L, L * M))
    for i in range(M - 1, -1, -1):
        e = -G[i] * e + x[i]
        for l in range(64):
            out[l // L, (l % L) * M + i] = e[l]
    return out


def sequential(km, a, d4, rbt, gg, bot):
    """The upstream order: elimination downward, substitution upward."""
    qp = np.zeros(km + 1)
    qp[1] = (d4[0] * a[0] + a[1]) * rbt[0]
    for k in range(2, km + 1):
        qp[k] = (3 * (a[k - 2] + d4[k - 1] * a[k - 1]) - qp[k - 1]) * rbt[k - 1]
    Q = np.zeros(km + 2)
    Q[km + 1] = (bot[0] * a[km - 1] + a[km - 2] - bot[1] * qp[km]) * bot[2]
    for k in range(km, 0, -1):
        Q[k] = qp[k] - gg[k - 1] * Q[k + 1]
    return Q


@pytest.mark.parametrize("M,L,km", [(2, 16, 8), (2, 16, 15), (2, 16, 16), (2, 16, 31), (4, 16, 32), (4, 16, 40),
                                    (4, 16, 62), (4, 16, 63), (4, 32, 64), (4, 32, 100), (4, 32, 125),
                                    (4, 32, 126), (4, 32, 127)])
def test_lane_scans_equal_sequential_solve(M, L, km):
    rng = np.random.default_rng(km * 7 + L)
    ncol, nl = 64 // L, L * M
    a = np.zeros((ncol, nl))
    d4, rbt, gg = np.zeros((ncol, nl)), np.zeros((ncol, nl)), np.zeros((ncol, nl))
    a[:, :km] = rng.uniform(-1, 1, (ncol, km))
    d4[:, :km] = rng.uniform(0.5, 2.0, (ncol, km))
    rbt[:, :km] = rng.uniform(0.2, 0.45, (ncol, km))
    gg[:, :km] = rng.uniform(0.1, 0.4, (ncol, km))
    bot = rng.uniform(0.5, 1.5, (ncol, 3))
    got = lane_solve(M, L, km, a, d4, rbt, gg, bot)
    for c in range(ncol):
        want = sequential(km, a[c], d4[c], rbt[c], gg[c], bot[c])
        np.testing.assert_allclose(got[c, :km + 1], want[1:km + 2], rtol=1e-12, atol=1e-12)


def test_a_wrong_carry_fails(monkeypatch):
    # the model detects a plausible mistake: slopes zeroed where a lane has no
    # partner (the first form of the kernel's 32-lane scan)
    import sys
    rng = np.random.default_rng(3)
    M, L, km = 4, 32, 127
    a = rng.uniform(-1, 1, (2, 128))
    d4, rbt, gg = rng.uniform(0.5, 2, (2, 128)), rng.uniform(0.2, 0.45, (2, 128)), rng.uniform(0.1, 0.4, (2, 128))
    for arr in (a, d4, rbt, gg):
        arr[:, km:] = 0
    bot = rng.uniform(0.5, 1.5, (2, 3))
    want = sequential(km, a[0], d4[0], rbt[0], gg[0], bot[0])
    good = lane_solve(M, L, km, a, d4, rbt, gg, bot)
    assert np.abs(good[0, :km + 1] - want[1:]).max() < 1e-12
    keep_old = dpp

    def zeroing(v, ctrl, old, bound_zero, row_mask=0xF):
        if old == 1.0 and not bound_zero:  # a slope move: read 0 instead of keeping the product
            return keep_old(v, ctrl, 0.0, True, row_mask)
        return keep_old(v, ctrl, old, bound_zero, row_mask)

    monkeypatch.setattr(sys.modules[__name__], "dpp", zeroing)
    bad = lane_solve(M, L, km, a, d4, rbt, gg, bot)
    assert np.abs(bad[0, :km + 1] - want[1:]).max() > 1e-6
