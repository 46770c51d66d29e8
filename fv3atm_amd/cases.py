"""Deterministic synthetic inputs for the tracer transport + remap path.

Python face of ``csrc/test_cases.c`` (analytic gnomonic equal-distance
cubed-sphere grid with neighbour-filled halos, stream-function winds, tracer
patterns, perturbed layer thicknesses).  Upstream these arrays come from the
dycore's grid generator and ``dyn_core`` (``cx, cy, mfx, mfy, dp1``) and from
the restart state (``q, delp, pt``); here they are generated.

Every array is a numpy float64 array in Fortran order with the dycore's
bounds (``isd:ied, jsd:jed`` with a 3-cell halo, ``1:npz``, ``1:nq``), as at
ccpp/data/CCPP_typedefs.F90:1580-1590.  Multi-tile arrays carry the tile as
the last (outermost) dimension.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field

import numpy as np

NG = 3
RADIUS = 6.3712e6
PTOP = 100.0
PS = 1.0e5

_here = os.path.dirname(os.path.abspath(__file__))
_lib = None


def _cases_lib():
    global _lib
    if _lib is None:
        path = os.path.join(_here, "libfv3_cases.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        P = ctypes.c_void_p
        i, d = ctypes.c_int, ctypes.c_double
        lib.fv3c_grid.argtypes = [i, i, d] + [P] * 11
        lib.fv3c_winds.argtypes = [i, i, i, i, P, d, P, d, P, P, P, P, P, P, P, P, P, P]
        lib.fv3c_tracer.argtypes = [i, i, i, P, P]
        lib.fv3c_delp.argtypes = [i, i, i, i, P, ctypes.c_uint64, P, d, P]
        lib.fv3c_halo_map.argtypes = [i, i, P]
        lib.fv3c_cell_centers.argtypes = [i, i, P]
        lib.fv3c_uniform.argtypes = [ctypes.c_uint64, i, i, i, i, i]
        lib.fv3c_uniform.restype = d
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@dataclass
class Grid:
    """fv_grid_type subset for one tile (moving_nest/fv_moving_nest.F90:81)."""
    n: int
    tile: int
    area: np.ndarray
    rarea: np.ndarray
    dx: np.ndarray
    dy: np.ndarray
    dxa: np.ndarray
    dya: np.ndarray
    sin_sg: np.ndarray
    del6_u: np.ndarray
    del6_v: np.ndarray
    xyz_c: np.ndarray
    da_min: float = 0.0
    sw_corner: bool = True
    se_corner: bool = True
    nw_corner: bool = True
    ne_corner: bool = True
    bounded_domain: bool = False
    grid_type: int = 0
    xyz_n: np.ndarray = None  # corner unit vectors (3, isd:ied+1, jsd:jed+1), halo-folded

    @property
    def npx(self):
        return self.n + 1

    @property
    def npy(self):
        return self.n + 1


def make_grid(n: int, tile: int = 1, radius: float = RADIUS) -> Grid:
    lib = _cases_lib()
    nd = n + 2 * NG
    z = lambda *s: np.zeros(s, dtype=np.float64, order="F")
    g = dict(area=z(nd, nd), rarea=z(nd, nd), dx=z(nd, nd + 1), dy=z(nd + 1, nd),
             dxa=z(nd, nd), dya=z(nd, nd), sin_sg=z(nd, nd, 9), del6_u=z(nd, nd + 1),
             del6_v=z(nd + 1, nd), xyz_c=z(3, nd, nd), xyz_n=z(3, nd + 1, nd + 1))
    rc = lib.fv3c_grid(n, tile, radius, _p(g["area"]), _p(g["rarea"]), _p(g["dx"]), _p(g["dy"]),
                       _p(g["dxa"]), _p(g["dya"]), _p(g["sin_sg"]), _p(g["del6_u"]),
                       _p(g["del6_v"]), _p(g["xyz_c"]), _p(g["xyz_n"]))
    if rc:
        raise ValueError(f"fv3c_grid failed for n={n} tile={tile}")
    return Grid(n=n, tile=tile, **g)


def sphere_da_min(grids) -> float:
    return float(min(g.area[NG:-NG, NG:-NG].min() for g in grids))


def halo_map(n: int, tile: int) -> np.ndarray:
    """Flat source index (into a 6-tile stack of (isd:ied, jsd:jed) slabs) of
    every edge-halo cell of ``tile``; -1 elsewhere."""
    nd = n + 2 * NG
    m = np.zeros((nd, nd), dtype=np.int64, order="F")
    _cases_lib().fv3c_halo_map(n, tile, _p(m))
    return m


def uniform(seed, tile, i, j, k, fld) -> float:
    return _cases_lib().fv3c_uniform(seed, tile, i, j, k, fld)


def hybrid_coefs(npz: int, ptop: float = PTOP):
    """Uniform-thickness hybrid-sigma ak/bk: pe(k) = ptop + (k-1)/npz (ps - ptop)."""
    bk = np.arange(npz + 1, dtype=np.float64) / npz
    ak = ptop * (1.0 - bk)
    return ak, bk


def winds(grid: Grid, npz: int, kind: int, par, speed_k, dt: float, dp_ref):
    lib = _cases_lib()
    n, nd = grid.n, grid.n + 2 * NG
    z = lambda *s: np.zeros(s, dtype=np.float64, order="F")
    cx, cy = z(n + 1, nd, npz), z(nd, n + 1, npz)
    mfx, mfy = z(n + 1, n, npz), z(n, n + 1, npz)
    par = np.ascontiguousarray(par, dtype=np.float64)
    speed_k = np.ascontiguousarray(speed_k, dtype=np.float64)
    dp_ref = np.ascontiguousarray(dp_ref, dtype=np.float64)
    lib.fv3c_winds(n, grid.tile, npz, kind, _p(par), RADIUS, _p(speed_k), dt, _p(dp_ref),
                   _p(grid.dxa), _p(grid.dy), _p(grid.dya), _p(grid.dx), _p(grid.sin_sg),
                   _p(cx), _p(cy), _p(mfx), _p(mfy))
    return cx, cy, mfx, mfy


def tracer2d(grid: Grid, kind: int, par) -> np.ndarray:
    nd = grid.n + 2 * NG
    q = np.zeros((nd, nd), dtype=np.float64, order="F")
    par = np.ascontiguousarray(par, dtype=np.float64)
    _cases_lib().fv3c_tracer(grid.n, grid.tile, kind, _p(par), _p(q))
    return q


def delp_field(grid: Grid, npz: int, kind: int, amp, seed: int, dp_ref,
               ps_minus_ptop: float = PS - PTOP) -> np.ndarray:
    nd = grid.n + 2 * NG
    d = np.zeros((nd, nd, npz), dtype=np.float64, order="F")
    amp = np.ascontiguousarray(amp, dtype=np.float64)
    dp_ref = np.ascontiguousarray(dp_ref, dtype=np.float64)
    _cases_lib().fv3c_delp(grid.n, grid.tile, npz, kind, _p(amp), seed, _p(dp_ref),
                           ps_minus_ptop, _p(d))
    return d


def std_height(p):
    """Standard-atmosphere height (m) of pressure p (Pa), troposphere lapse 6.5 K/km."""
    return 288.15 / 6.5e-3 * (1.0 - (np.asarray(p) / 1.01325e5) ** (287.04 * 6.5e-3 / 9.80665))


@dataclass
class Case:
    """One synthetic instance of the path's inputs (single tile or sphere)."""
    name: str
    n: int
    npz: int
    nq: int
    ntiles: int
    grids: list
    q: np.ndarray          # (nd, nd, npz, nq[, ntiles])
    dp1: np.ndarray        # (nd, nd, npz[, ntiles])
    cx: np.ndarray
    cy: np.ndarray
    mfx: np.ndarray
    mfy: np.ndarray
    ak: np.ndarray
    bk: np.ndarray
    ptop: float
    pt: np.ndarray | None
    hord: int = 8
    kord_tr: int = 9
    kord_tm: int = 10
    q_split: int = 0
    dt: float = 1.0
    lim_fac: float = 1.0
    meta: dict = field(default_factory=dict)


def _tracer_params(nq: int, seed: int, tile: int):
    """Families cycle bell / slotted cylinder / smooth with seeded centres
    near the centre of ``tile``."""
    from math import atan2, asin
    g = make_grid(8, tile)
    c = g.xyz_c[:, NG + 4, NG + 4]
    lon0, lat0 = atan2(c[1], c[0]), asin(c[2])
    out = []
    r0 = (math.pi / 2.0) / 3.0
    for iq in range(nq):
        fam = iq % 3
        dl = 0.35 * uniform(seed, tile, iq, 0, 0, 11)
        dphi = 0.35 * uniform(seed, tile, iq, 1, 0, 11)
        if fam == 0:
            out.append((0, [lon0 + dl, lat0 + dphi, r0, 1.0e-6]))
        elif fam == 1:
            out.append((1, [lon0 + dl, lat0 + dphi, r0]))
        else:
            out.append((2, []))
    return out


def make_case(name: str, n: int, npz: int, nq: int, *, ntiles: int = 1, tile: int = 1,
              wind: str = "solid", courant: float = 0.4, jet: bool = False,
              delp_kind: int = 0, delp_amp=(0.1,), seed: int = 0, q_split: int = 0,
              with_pt: bool = False, tilt: float = math.pi / 4, tracers=None,
              build_q: bool = True) -> Case:
    """Build a case.  ``courant`` is the target max |Courant| over the tile
    interior (for ``jet`` the max over levels)."""
    tiles = list(range(1, 7)) if ntiles == 6 else [tile]
    grids = [make_grid(n, t) for t in tiles]
    da_min = sphere_da_min([make_grid(n, t) for t in range(1, 7)]) if ntiles == 6 else \
        float(grids[0].area[NG:-NG, NG:-NG].min())
    for g in grids:
        g.da_min = da_min
    ak, bk = hybrid_coefs(npz)
    dp_ref = np.full(npz, (PS - PTOP) / npz)
    pmid = PTOP + (np.arange(npz) + 0.5) * (PS - PTOP) / npz
    if jet:
        z = std_height(pmid)
        speed = 10.0 + 60.0 * np.exp(-((z - 10.0e3) / 3.0e3) ** 2)
    else:
        speed = np.ones(npz)
    kind, par = (0, [tilt]) if wind == "solid" else (1, [0.25, 12.0 * 86400.0])
    # unit-dt pass to scale the time step to the requested Courant number
    cmax = 0.0
    for g in grids:
        cx1, cy1, _, _ = winds(g, npz, kind, par, speed, 1.0, dp_ref)
        cmax = max(cmax, np.abs(cx1[:, NG:-NG, :]).max(), np.abs(cy1[NG:-NG, :, :]).max())
    dt = courant / cmax
    parts = [winds(g, npz, kind, par, speed, dt, dp_ref) for g in grids]
    st = lambda i: np.stack([p[i] for p in parts], axis=-1) if ntiles > 1 else parts[0][i]
    cx, cy, mfx, mfy = st(0), st(1), st(2), st(3)
    dp = [delp_field(g, npz, delp_kind, list(delp_amp), seed, dp_ref) for g in grids]
    dp1 = np.stack(dp, axis=-1) if ntiles > 1 else dp[0]
    if tracers is None:
        tracers = _tracer_params(nq, seed, tiles[0] if ntiles == 1 else 1)
    nd = n + 2 * NG
    lev = 1.0 + 0.5 * np.arange(npz) / max(npz, 1)
    bases = [[tracer2d(g, knd, pr) for (knd, pr) in tracers] for g in grids]
    if build_q:
        q = np.zeros((nd, nd, npz, nq, len(grids)), dtype=np.float64, order="F")
        for t in range(len(grids)):
            for iq in range(nq):
                q[:, :, :, iq, t] = bases[t][iq][:, :, None] * lev[None, None, :]
        if ntiles == 1:
            q = np.asfortranarray(q[..., 0])
    else:
        q = None  # q(i,j,k,iq) = bases[tile][iq](i,j) * lev(k); assemble where it is used
    pt = None
    if with_pt:
        pmid3 = PTOP + np.cumsum(dp1, axis=2) - 0.5 * dp1
        # smooth profile from 184 K at the model top to 300 K at the surface
        # (a floor clamp would put a kink in every column)
        pt = np.asfortranarray(184.0 + 116.0 * (pmid3 / PS) ** (287.04 * 6.5e-3 / 9.80665))
    return Case(name=name, n=n, npz=npz, nq=nq, ntiles=ntiles, grids=grids, q=q, dp1=dp1,
                cx=cx, cy=cy, mfx=mfx, mfy=mfy, ak=ak, bk=bk, ptop=PTOP, pt=pt,
                q_split=q_split, dt=dt, meta=dict(speed=speed, wind=wind, courant=courant,
                                                  q_bases=bases, lev=lev))


def config(idx: int, **over) -> Case:
    """The workload shapes of the path's benchmark configurations.
    0: C48 L32 Q1 (one tile, solid body, Courant 0.4, kord 9 remap);
    1: C384 L64 Q10 (deformational, forced nsplit=2);
    2: C768 L127 Q30 (+pt) jet profile, per-level nsplit;
    3: C192 L64 Q4 (six forced sub-cycles);
    4: 6 x C384 L127 Q20 sphere."""
    if idx == 0:
        kw = dict(name="C48L32Q1", n=48, npz=32, nq=1, courant=0.4, delp_kind=0,
                  delp_amp=(0.1,), tracers=[(0, None)])
        kw.update(over)
        if kw["tracers"] == [(0, None)]:
            g = make_grid(8, 1)
            c = g.xyz_c[:, NG + 4, NG + 4]
            kw["tracers"] = [(0, [math.atan2(c[1], c[0]), math.asin(c[2]),
                                  (math.pi / 2.0) / 3.0, 1.0e-6])]
        return make_case(**kw)
    if idx == 1:
        kw = dict(name="C384L64Q10", n=384, npz=64, nq=10, wind="deform", courant=0.8,
                  q_split=2, delp_kind=1, delp_amp=(0.05, 0.1))
    elif idx == 2:
        kw = dict(name="C768L127Q30", n=768, npz=127, nq=30, jet=True, courant=2.5,
                  delp_kind=1, delp_amp=(0.05, 0.2), with_pt=True)
    elif idx == 3:
        kw = dict(name="C192L64Q4", n=192, npz=64, nq=4, courant=0.9, q_split=6,
                  delp_kind=1, delp_amp=(0.05, 0.1))
    elif idx == 4:
        kw = dict(name="6xC384L127Q20", n=384, npz=127, nq=20, ntiles=6, courant=0.9,
                  delp_kind=1, delp_amp=(0.05, 0.2))
    else:
        raise ValueError(idx)
    kw.update(over)
    return make_case(**kw)


RDGAS, CP_AIR, GRAV = 287.05, 1004.6, 9.80665
KAPPA = RDGAS / CP_AIR


def l2e_state(case: Case, seed: int = 0, hydrostatic: bool = False) -> dict:
    """The rest of the dynamical state that Lagrangian_to_Eulerian remaps, on
    the case's grid and Lagrangian thickness dp1: D-grid winds u (isd:ied,
    jsd:jed+1, km) and v (isd:ied+1, jsd:jed, km) from a sheared solid-body
    flow plus seeded waves, vertical velocity w (isd:ied, jsd:jed, km) and its
    surface value ws (is:ie, js:je), hydrostatic layer depths delz (is:ie,
    js:je, km) and pkz (is:ie, js:je, km), pt, and the tracers.  Arrays carry
    the tile as the last axis for six-tile cases."""
    n, km = case.n, case.npz
    nd = n + 2 * NG
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.5, 1.5, 6)
    tiles = range(case.ntiles)
    out = {k: [] for k in ("u", "v", "w", "ws", "delz", "pkz")}
    for t in tiles:
        g = case.grids[t]
        dp = case.dp1[..., t] if case.ntiles > 1 else case.dp1
        pe = PTOP + np.concatenate([np.zeros((nd, nd, 1)), np.cumsum(dp, axis=2)], axis=2)
        lev = np.arange(km)[None, None, :] / max(km - 1, 1)
        # winds: functions of the face-point positions (halo-folded), so the
        # halos are the neighbours' values in this face's directions only
        # approximately; Lagrangian_to_Eulerian reads the compute domain
        X = g.xyz_n
        yf = 0.5 * (X[:, 1:, :] + X[:, :-1, :])   # y-faces (nd, nd+1)
        xf = 0.5 * (X[:, :, 1:] + X[:, :, :-1])   # x-faces (nd+1, nd)
        su = 10.0 + 30.0 * np.sin(np.pi * lev)
        u = (amp[0] * np.sin(2.0 * yf[1] + 0.4) + amp[1] * yf[2])[..., None] * su
        v = (amp[2] * np.cos(1.5 * xf[0] - 0.2) + amp[3] * xf[1] * xf[2])[..., None] * su
        c = g.xyz_c
        w = (amp[4] * np.sin(3.0 * c[0] + 1.0) * np.cos(2.0 * c[1]))[..., None] * (0.2 + np.sin(np.pi * lev))
        ws = 0.05 * amp[5] * np.sin(2.0 * c[2] + 0.3)[NG:-NG, NG:-NG]
        # hydrostatic thickness and p^kappa layer means of the Lagrangian layers
        T = 300.0 - 70.0 * lev * 0 + 0.0 * dp  # placeholder replaced below
        lp = np.log(pe)
        pk = np.exp(KAPPA * lp)
        pkz = (pk[..., 1:] - pk[..., :-1]) / (KAPPA * (lp[..., 1:] - lp[..., :-1]))
        ptv = case.pt[..., t] if (case.pt is not None and case.ntiles > 1) else case.pt
        if ptv is None:
            ptv = 184.0 + 116.0 * ((pe[..., 1:] + pe[..., :-1]) / (2 * PS)) ** KAPPA
        T = ptv * pkz
        delz = -RDGAS / GRAV * T * (lp[..., 1:] - lp[..., :-1])
        out["u"].append(np.asfortranarray(u))
        out["v"].append(np.asfortranarray(v))
        out["w"].append(np.asfortranarray(w))
        out["ws"].append(np.asfortranarray(ws))
        out["delz"].append(np.asfortranarray(delz[NG:-NG, NG:-NG]))
        out["pkz"].append(np.asfortranarray(pkz[NG:-NG, NG:-NG]))
    st = {k: (np.stack(v, axis=-1) if case.ntiles > 1 else v[0]) for k, v in out.items()}
    st = {k: np.asfortranarray(v) for k, v in st.items()}
    st["delp"] = np.asfortranarray(case.dp1.copy())
    pt = case.pt
    if pt is None:
        lev_pe = PTOP + np.cumsum(case.dp1, axis=2) - 0.5 * case.dp1
        pt = 184.0 + 116.0 * (lev_pe / PS) ** KAPPA
    st["pt"] = np.asfortranarray(pt.copy())
    st["q"] = np.asfortranarray(case.q.copy())
    if hydrostatic:
        del st["w"], st["ws"], st["delz"]
    return st


def c_grid_winds(case, n_split: int, amp: float = 0.2):
    """C-grid winds of n_split acoustic sub-steps whose Courant numbers over
    the step add up to about the case's cx, cy: the wind the case's Courant
    numbers imply (uc = cx dxa(upwind) / dt) modulated per sub-step and
    level by 1 + amp sin(...).  Returns uc (n+1, nd, npz, ntiles, n_split),
    vc (nd, n+1, npz, ntiles, n_split), Fortran order."""
    n, nd, npz, nt = case.n, case.n + 2 * NG, case.npz, case.ntiles
    cx = case.cx.reshape(n + 1, nd, npz, nt, order="F")
    cy = case.cy.reshape(nd, n + 1, npz, nt, order="F")
    uc = np.zeros((n + 1, nd, npz, nt, n_split), order="F")
    vc = np.zeros((nd, n + 1, npz, nt, n_split), order="F")
    for t, g in enumerate(case.grids):
        dxa_m, dxa_p = g.dxa[NG - 1:NG + n, :], g.dxa[NG:NG + n + 1, :]       # cells i-1, i of face i
        dya_m, dya_p = g.dya[:, NG - 1:NG + n], g.dya[:, NG:NG + n + 1]
        u0 = np.where(cx[..., t] > 0, cx[..., t] * dxa_m[..., None], cx[..., t] * dxa_p[..., None]) / case.dt
        v0 = np.where(cy[..., t] > 0, cy[..., t] * dya_m[..., None], cy[..., t] * dya_p[..., None]) / case.dt
        for it in range(n_split):
            ph = 2.0 * np.pi * (it + 0.5) / n_split + 0.37 * np.arange(npz)
            uc[..., t, it] = u0 * (1.0 + amp * np.sin(ph))
            vc[..., t, it] = v0 * (1.0 + amp * np.cos(ph))
    return uc, vc
