"""fv3atm_amd -- MI355X-native FV3 tracer transport and vertical remap.

Host-side mirror of the upstream Fortran interfaces of the path
(GFDL_atmos_cubed_sphere: tp_core.F90 fv_tp_2d, fv_tracer2d.F90
tracer_2d / tracer_2d_1L, fv_mapz.F90 map1_ppm / map_scalar / mapn_tracer /
cs_profile / Lagrangian_to_Eulerian), calling the HIP kernels of
libfv3atm_amd.so through its C-ABI (include/fv3atm_amd.h).  Same argument
names and meanings; numpy arrays in the dycore's Fortran layout; failures
raise RuntimeError, as mpp_error(FATAL) aborts upstream.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _ffi
from ._ffi import (AGRID, CGRID, DGRID, FORM_FUSED32, FORM_FUSED64, FORM_PLAIN64,
                   NOALLOC_SCRATCH_BYTES, check, ptr)

NG = 3


def _f64(a, name):
    if a is None:
        return None
    if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags.f_contiguous):
        raise TypeError(f"{name}: a Fortran-contiguous float64 array is required")
    return a


class Context:
    """Device-resident domain: bounds, levels, tiles and grid metrics
    (the gridstruct / bd / domain handles tracer_2d receives upstream)."""

    def __init__(self, grids, npz: int, bd=None):
        grids = list(grids)
        n = grids[0].n
        if bd is None:
            bd = _ffi.Bounds(1, n, 1, n, 1 - NG, n + NG, 1 - NG, n + NG)
        self.bd = bd
        self.n = n
        self.npx = self.npy = n + 1
        self.npz = npz
        self.ntiles = len(grids)
        self.grids = grids
        h = ctypes.c_void_p()
        check(_ffi.lib().fv3_ctx_create(ctypes.byref(bd), self.npx, self.npy, npz, self.ntiles,
                                        ctypes.byref(h)), "fv3_ctx_create")
        self._h = h
        for t, g in enumerate(grids, start=1):
            gs = _ffi.Grid(*[ptr(_f64(getattr(g, k), k)) for k in
                             ("area", "rarea", "dx", "dy", "dxa", "dya", "sin_sg", "del6_u",
                              "del6_v")],
                           g.da_min, int(g.sw_corner), int(g.se_corner), int(g.nw_corner),
                           int(g.ne_corner), int(g.bounded_domain), g.grid_type)
            check(_ffi.lib().fv3_ctx_set_grid(h, t, ctypes.byref(gs)), "fv3_ctx_set_grid")

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            _ffi.lib().fv3_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fv_tp_2d(ctx: Context, q, crx, cry, npx, npy, hord, fx, fy, xfx, yfx, ra_x, ra_y,
             lim_fac=1.0, mfx=None, mfy=None, mass=None, nord=None, damp_c=None):
    """tp_core.F90 fv_tp_2d: fills fx, fy; modifies q's corner halos."""
    n = ctypes.c_int(nord) if nord is not None else None
    dc = ctypes.c_double(damp_c) if damp_c is not None else None
    check(_ffi.lib().fv3_fv_tp_2d(
        ctx.handle, ptr(_f64(q, "q")), ptr(_f64(crx, "crx")), ptr(_f64(cry, "cry")), npx, npy,
        hord, ptr(_f64(fx, "fx")), ptr(_f64(fy, "fy")), ptr(_f64(xfx, "xfx")),
        ptr(_f64(yfx, "yfx")), ptr(_f64(ra_x, "ra_x")), ptr(_f64(ra_y, "ra_y")), lim_fac,
        ptr(_f64(mfx, "mfx")), ptr(_f64(mfy, "mfy")), ptr(_f64(mass, "mass")),
        ctypes.byref(n) if n is not None else None, ctypes.byref(dc) if dc is not None else None),
        "fv_tp_2d")


def _tracer(fn, name, ctx, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split, dt,
            id_divg, nord_tr, trdm, lim_fac):
    nsplit = np.zeros(npz, dtype=np.int32)
    check(fn(ctx.handle, ptr(_f64(q, "q")), ptr(_f64(dp1, "dp1")), ptr(_f64(mfx, "mfx")),
             ptr(_f64(mfy, "mfy")), ptr(_f64(cx, "cx")), ptr(_f64(cy, "cy")), npx, npy, npz, nq,
             hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac, ptr(nsplit)), name)
    return nsplit


def tracer_2d_1L(ctx: Context, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split=0,
                 dt=1.0, id_divg=0, nord_tr=0, trdm=0.0, lim_fac=1.0):
    """fv_tracer2d.F90 tracer_2d_1L (in place); returns nsplit(k)."""
    return _tracer(_ffi.lib().fv3_tracer_2d_1l, "tracer_2d_1L", ctx, q, dp1, mfx, mfy, cx, cy,
                   npx, npy, npz, nq, hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac)


def tracer_2d(ctx: Context, q, dp1, mfx, mfy, cx, cy, npx, npy, npz, nq, hord, q_split=0,
              dt=1.0, id_divg=0, nord_tr=0, trdm=0.0, lim_fac=1.0):
    """fv_tracer2d.F90 tracer_2d (in place); returns nsplit(k)."""
    return _tracer(_ffi.lib().fv3_tracer_2d, "tracer_2d", ctx, q, dp1, mfx, mfy, cx, cy, npx, npy,
                   npz, nq, hord, q_split, dt, id_divg, nord_tr, trdm, lim_fac)


def map1_ppm(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord):
    """fv_mapz.F90 map1_ppm (row j)."""
    check(_ffi.lib().fv3_map1_ppm(km, ptr(_f64(pe1, "pe1")), ptr(_f64(q1, "q1")),
                                  ptr(_f64(qs, "qs")), kn, ptr(_f64(pe2, "pe2")),
                                  ptr(_f64(q2, "q2")), i1, i2, j, ibeg, iend, jbeg, jend, iv,
                                  kord), "map1_ppm")


def map_scalar(km, pe1, q1, qs, kn, pe2, q2, i1, i2, j, ibeg, iend, jbeg, jend, iv, kord, q_min):
    """fv_mapz.F90 map_scalar (row j)."""
    check(_ffi.lib().fv3_map_scalar(km, ptr(_f64(pe1, "pe1")), ptr(_f64(q1, "q1")),
                                    ptr(_f64(qs, "qs")), kn, ptr(_f64(pe2, "pe2")),
                                    ptr(_f64(q2, "q2")), i1, i2, j, ibeg, iend, jbeg, jend, iv,
                                    kord, q_min), "map_scalar")


def mapn_tracer(nq, km, pe1, pe2, q1, dp2, kord, j, i1, i2, isd, ied, jsd, jed, q_min, fill):
    """fv_mapz.F90 mapn_tracer (row j, q1 in place)."""
    kord = np.ascontiguousarray(kord, dtype=np.int32)
    check(_ffi.lib().fv3_mapn_tracer(nq, km, ptr(_f64(pe1, "pe1")), ptr(_f64(pe2, "pe2")),
                                     ptr(_f64(q1, "q1")), ptr(_f64(dp2, "dp2")), ptr(kord), j,
                                     i1, i2, isd, ied, jsd, jed, q_min, int(fill)), "mapn_tracer")


def cs_profile(qs, a4, delp, km, i1, i2, iv, kord):
    """fv_mapz.F90 cs_profile (a4 in place)."""
    check(_ffi.lib().fv3_cs_profile(ptr(_f64(qs, "qs")), ptr(_f64(a4, "a4")),
                                    ptr(_f64(delp, "delp")), km, i1, i2, iv, kord), "cs_profile")


def lagrangian_to_eulerian_tracers(ctx: Context, km, nq, ak, bk, ptop, delp, q, pt, kord_tr,
                                   kord_tm, t_min=184.0, fill=False):
    """Tracer/pt subset of fv_mapz.F90 Lagrangian_to_Eulerian (in place)."""
    kord_tr = np.ascontiguousarray(np.broadcast_to(kord_tr, (nq,)), dtype=np.int32)
    check(_ffi.lib().fv3_lagrangian_to_eulerian_tracers(
        ctx.handle, km, nq, ptr(np.ascontiguousarray(ak, dtype=np.float64)),
        ptr(np.ascontiguousarray(bk, dtype=np.float64)), ptop, ptr(_f64(delp, "delp")),
        ptr(_f64(q, "q")), ptr(_f64(pt, "pt")), ptr(kord_tr), kord_tm, t_min, int(fill)),
        "Lagrangian_to_Eulerian")


def _sizes(ctx: Context, nq: int):
    """Element counts of the device arrays of a step on ctx (every tile)."""
    b, nt, npz = ctx.bd, ctx.ntiles, ctx.npz
    nid, njd = b.ied - b.isd + 1, b.jed - b.jsd + 1
    nix, niy, nj = b.ie - b.is_ + 2, b.ie - b.is_ + 1, b.je - b.js + 1
    lev = npz * nt
    return dict(q=nid * njd * npz * nq * nt, pt=nid * njd * lev, dp=nid * njd * lev,
                cx=nix * njd * lev, cy=nid * (nj + 1) * lev, mfx=nix * nj * lev,
                mfy=niy * (nj + 1) * lev, ak=npz + 1)


def _dev_args(ctx: Context, nq, spec):
    """Checks torch device arguments before their addresses cross the C-ABI:
    on the context's device, contiguous, of the element type the form
    computes in, and at least as long as the step reads or writes.
    spec: (name, tensor or None, dtype, size key, required)."""
    import torch
    n = _sizes(ctx, nq)
    dev = None
    for name, t, dt, key, required in spec:
        if t is None:
            if required:
                raise TypeError(f"{name}: a device tensor is required")
            continue
        if not isinstance(t, torch.Tensor) or not t.is_cuda:
            raise TypeError(f"{name}: a torch tensor on the HIP device is required")
        if t.dtype != dt:
            raise TypeError(f"{name}: dtype {t.dtype}, expected {dt}")
        if not t.is_contiguous():
            raise ValueError(f"{name}: a contiguous tensor is required")
        if t.numel() < n[key]:
            raise ValueError(f"{name}: {t.numel()} elements, the step needs {n[key]}")
        if dev is None:
            dev = t.device
        elif t.device != dev:
            raise ValueError(f"{name}: on {t.device}, the other arguments on {dev}")


def _stream(stream):
    """The caller's stream, or torch's current stream on the current device."""
    if stream is not None:
        return stream
    import torch
    return torch.cuda.current_stream().cuda_stream


def _step_spec(form, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk):
    import torch
    if form not in (FORM_PLAIN64, FORM_FUSED64, FORM_FUSED32):
        raise ValueError(f"unknown form {form}")
    w = torch.float32 if form == FORM_FUSED32 else torch.float64
    f64 = torch.float64
    return [("q_in", q_in, w, "q", True), ("q_out", q_out, w, "q", True), ("pt", pt, w, "pt", False),
            ("dp1", dp1, f64, "dp", True), ("delp_out", delp_out, f64, "dp", True),
            ("cx", cx, w, "cx", True), ("cy", cy, w, "cy", True), ("mfx", mfx, w, "mfx", True),
            ("mfy", mfy, w, "mfy", True), ("ak", ak, f64, "ak", True), ("bk", bk, f64, "ak", True)]


def step_dev(ctx: Context, form, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk,
             ptop, hord=8, kord_tr=9, kord_tm=10, t_min=184.0, nsplit_cap=0, stream=None):
    """Device-resident advection + remap step on torch device tensors (on
    ``stream``, a hipStream_t handle; default: torch's current stream)."""
    _dev_args(ctx, nq, _step_spec(form, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk))
    check(_ffi.lib().fv3_step_dev(ctx.handle, form, nq, ptr(q_in), ptr(q_out), ptr(pt), ptr(dp1),
                                  ptr(delp_out), ptr(cx), ptr(cy), ptr(mfx), ptr(mfy), ptr(ak),
                                  ptr(bk), ptop, hord, kord_tr, kord_tm, t_min, nsplit_cap,
                                  _stream(stream)), "fv3_step_dev")


def step_dev_noalloc(ctx: Context, nq, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy, ak, bk,
                     ptop, hord=8, kord_tr=9, kord_tm=10, t_min=184.0, nsplit_cap=0, stream=None):
    """The FP32 step of step_dev with its scratch in the library's static
    per-device array (no device allocation; at most NOALLOC_SCRATCH_BYTES)."""
    _dev_args(ctx, nq, _step_spec(FORM_FUSED32, q_in, q_out, pt, dp1, delp_out, cx, cy, mfx, mfy,
                                  ak, bk))
    check(_ffi.lib().fv3_step_dev_noalloc(ctx.handle, FORM_FUSED32, nq, ptr(q_in), ptr(q_out),
                                          ptr(pt), ptr(dp1), ptr(delp_out), ptr(cx), ptr(cy),
                                          ptr(mfx), ptr(mfy), ptr(ak), ptr(bk), ptop, hord, kord_tr,
                                          kord_tm, t_min, nsplit_cap, _stream(stream)),
          "fv3_step_dev_noalloc")


def _tlad_spec(names_tensors, with_remap):
    import torch
    f64 = torch.float64
    keys = dict(q0="q", dq0="q", q1="q", dq1="q", qbar1="q", qbar0="q", dp1="dp", delp_out="dp",
                cx="cx", cy="cy", mfx="mfx", mfy="mfy", ak="ak", bk="ak")
    return [(k, t, f64, keys[k], k != "delp_out" and (with_remap or k not in ("ak", "bk")))
            for k, t in names_tensors]


def step_tl_dev(ctx: Context, per_level, q_split, nq, q0, dq0, q1, dq1, dp1, delp_out, cx, cy, mfx,
                mfy, ak, bk, ptop, hord=8, kord_tr=9, stream=None):
    """Tangent-linear of the FP64 advection + tracer remap step (frozen
    branches) on torch device tensors: (q0, dq0) -> (q1, dq1)."""
    _dev_args(ctx, nq, _tlad_spec([("q0", q0), ("dq0", dq0), ("q1", q1), ("dq1", dq1), ("dp1", dp1),
                                   ("delp_out", delp_out), ("cx", cx), ("cy", cy), ("mfx", mfx),
                                   ("mfy", mfy), ("ak", ak), ("bk", bk)], kord_tr != 0))
    check(_ffi.lib().fv3_step_tl_dev(ctx.handle, int(per_level), q_split, nq, ptr(q0), ptr(dq0),
                                     ptr(q1), ptr(dq1), ptr(dp1), ptr(delp_out), ptr(cx), ptr(cy),
                                     ptr(mfx), ptr(mfy), ptr(ak), ptr(bk), ptop, hord, kord_tr,
                                     _stream(stream)), "fv3_step_tl_dev")


def step_ad_dev(ctx: Context, per_level, q_split, nq, q0, qbar1, qbar0, dp1, cx, cy, mfx, mfy, ak,
                bk, ptop, hord=8, kord_tr=9, stream=None):
    """Adjoint of the same step: qbar1 (adjoint of the output) -> qbar0."""
    _dev_args(ctx, nq, _tlad_spec([("q0", q0), ("qbar1", qbar1), ("qbar0", qbar0), ("dp1", dp1),
                                   ("cx", cx), ("cy", cy), ("mfx", mfx), ("mfy", mfy), ("ak", ak),
                                   ("bk", bk)], kord_tr != 0))
    check(_ffi.lib().fv3_step_ad_dev(ctx.handle, int(per_level), q_split, nq, ptr(q0), ptr(qbar1),
                                     ptr(qbar0), ptr(dp1), ptr(cx), ptr(cy), ptr(mfx), ptr(mfy),
                                     ptr(ak), ptr(bk), ptop, hord, kord_tr, _stream(stream)),
          "fv3_step_ad_dev")


def halo_update(ctx: Context, field, nslab, stream=None):
    """mpp_update_domains of a scalar on a six-tile context: the 3-cell edge
    halos of nslab (isd:ied, jsd:jed) slabs per tile (tiles outermost) from
    the neighbouring faces, on the device."""
    import torch
    if not (isinstance(field, torch.Tensor) and field.is_cuda and field.is_contiguous()):
        raise TypeError("field: a contiguous torch tensor on the HIP device is required")
    if field.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"field: float32 or float64 required, got {field.dtype}")
    b = ctx.bd
    need = 6 * nslab * (b.ied - b.isd + 1) * (b.jed - b.jsd + 1)
    if ctx.ntiles != 6 or field.numel() < need:
        raise ValueError(f"field: a six-tile context and {need} elements are required")
    check(_ffi.lib().fv3_halo_update(ctx.handle, ptr(field), nslab, field.element_size(),
                                     _stream(stream)), "fv3_halo_update")


def halo_update_vector(ctx: Context, u, v, nslab, grid, stream=None):
    """mpp_update_domains of a vector pair (grid AGRID, CGRID or DGRID: the
    staggering of u and v) on a six-tile context, components turned into
    each face's index directions across rotated cube edges."""
    import torch
    b = ctx.bd
    nid, njd = b.ied - b.isd + 1, b.jed - b.jsd + 1
    shapes = {_ffi.AGRID: (nid * njd, nid * njd), _ffi.CGRID: ((nid + 1) * njd, nid * (njd + 1)),
              _ffi.DGRID: (nid * (njd + 1), (nid + 1) * njd)}
    if grid not in shapes:
        raise ValueError(f"grid must be AGRID, CGRID or DGRID, got {grid}")
    for name, t, per in (("u", u, shapes[grid][0]), ("v", v, shapes[grid][1])):
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
            raise TypeError(f"{name}: a contiguous torch tensor on the HIP device is required")
        if t.dtype != u.dtype or t.dtype not in (torch.float32, torch.float64):
            raise TypeError(f"{name}: float32 or float64, the same for u and v")
        if ctx.ntiles != 6 or t.numel() < 6 * nslab * per:
            raise ValueError(f"{name}: a six-tile context and {6 * nslab * per} elements are required")
    check(_ffi.lib().fv3_halo_update_vector(ctx.handle, ptr(u), ptr(v), nslab, grid, u.element_size(),
                                            _stream(stream)), "fv3_halo_update_vector")


def _dev_field(name, t, need):
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise TypeError(f"{name}: a contiguous torch tensor on the HIP device is required")
    if t.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"{name}: float32 or float64 required, got {t.dtype}")
    if t.numel() < need:
        raise ValueError(f"{name}: {need} elements required, got {t.numel()}")


def _slab(ctx: Context):
    b = ctx.bd
    return (b.ied - b.isd + 1) * (b.jed - b.jsd + 1)


def field_checksum(ctx: Context, field, nslab, stream=None):
    """FMS mpp_chksum of each tile's compute domain of a device field of
    nslab (isd:ied, jsd:jed) slabs per tile: the sum modulo 2**64 of the
    values' bit patterns as two's-complement integers.  Returns ntiles ints."""
    _dev_field("field", field, ctx.ntiles * nslab * _slab(ctx))
    out = (ctypes.c_ulonglong * ctx.ntiles)()
    check(_ffi.lib().fv3_field_checksum(ctx.handle, ptr(field), nslab, field.element_size(), out,
                                        _stream(stream)), "fv3_field_checksum")
    return [int(v) for v in out]


def dyn_core_fluxes(ctx: Context, form, n_split, bdt, hord_dp, uc, vc, delp, dp1, cx, cy, mfx, mfy,
                    stream=None):
    """dyn_core's accumulation of the tracer inputs over n_split acoustic
    sub-steps of length bdt (see fv3_dyn_core_fluxes).  Device tensors of
    the form's element type: uc, vc the C-grid winds of every sub-step;
    delp in/out; dp1, cx, cy, mfx, mfy out."""
    import torch
    b = ctx.bd
    nid, njd = b.ied - b.isd + 1, b.jed - b.jsd + 1
    ni, nj = b.ie - b.is_ + 1, b.je - b.js + 1
    per = ctx.ntiles * ctx.npz
    dt = torch.float32 if form == FORM_FUSED32 else torch.float64
    need = dict(uc=n_split * per * (ni + 1) * njd, vc=n_split * per * nid * (nj + 1), delp=per * nid * njd,
                dp1=per * nid * njd, cx=per * (ni + 1) * njd, cy=per * nid * (nj + 1), mfx=per * (ni + 1) * nj,
                mfy=per * ni * (nj + 1))
    ts = dict(uc=uc, vc=vc, delp=delp, dp1=dp1, cx=cx, cy=cy, mfx=mfx, mfy=mfy)
    for k, t in ts.items():
        _dev_field(k, t, need[k])
        if t.dtype != dt:
            raise TypeError(f"{k}: {dt} required for form {form}, got {t.dtype}")
    check(_ffi.lib().fv3_dyn_core_fluxes(ctx.handle, form, n_split, bdt, hord_dp, ptr(uc), ptr(vc), ptr(delp),
                                         ptr(dp1), ptr(cx), ptr(cy), ptr(mfx), ptr(mfy), _stream(stream)),
          "fv3_dyn_core_fluxes")


def global_mass(ctx: Context, q, nq, delp, stream=None):
    """Per-tracer global mass sum(q * delp * area) over every tile, level and
    compute-domain cell, FP64 with Neumaier compensation in a fixed order."""
    import torch
    per = ctx.ntiles * ctx.npz * _slab(ctx)
    _dev_field("q", q, per * nq)
    _dev_field("delp", delp, per)
    if delp.dtype != torch.float64:
        raise TypeError("delp: float64 required")
    out = np.zeros(nq, dtype=np.float64)
    check(_ffi.lib().fv3_global_mass(ctx.handle, ptr(q), nq, ptr(delp), q.element_size(),
                                     out.ctypes.data_as(_ffi.DP), _stream(stream)), "fv3_global_mass")
    return out


def _names(names):
    arr = (ctypes.c_char_p * len(names))(*[n.encode() for n in names])
    return arr


def tracer_restart_write(ctx: Context, prefix, names, q, tile_base=1, stream=None):
    """Writes <prefix>.tile<t>.nc (fv_tracer.res layout, one file per tile)
    from the device tracers q ((isd:ied, jsd:jed, npz, nq) per tile)."""
    _dev_field("q", q, ctx.ntiles * ctx.npz * len(names) * _slab(ctx))
    check(_ffi.lib().fv3_tracer_restart_write(ctx.handle, prefix.encode(), tile_base, len(names), _names(names),
                                              ptr(q), q.element_size(), _stream(stream)),
          "fv3_tracer_restart_write")


def tracer_restart_read(ctx: Context, prefix, names, q, tile_base=1, ignore_cksum=False, stream=None):
    """Reads the tracers named in names from <prefix>.tile<t>.nc into the
    compute domain of the device array q, checking each FMS checksum unless
    ignore_cksum."""
    _dev_field("q", q, ctx.ntiles * ctx.npz * len(names) * _slab(ctx))
    check(_ffi.lib().fv3_tracer_restart_read(ctx.handle, prefix.encode(), tile_base, len(names), _names(names),
                                             ptr(q), q.element_size(), int(bool(ignore_cksum)), _stream(stream)),
          "fv3_tracer_restart_read")


def history_write(ctx: Context, path, fields, pfull, phalf, *, out_bytes=4, quantize_nsd=0, tile_base=1,
                  lon=None, lat=None, ak=None, bk=None, time=0.0, time_units="hours since 1970-01-01 00:00:00",
                  time_iso="1970-01-01T00:00:00Z", stream=None):
    """Writes the history bundle file path.  fields: (name, long_name, units,
    device tensor, nlev) with nlev 1 or npz; pfull, phalf: host float arrays;
    lon, lat: host float64 (tile, j, i) arrays of the compute domain or None."""
    import numpy as np
    fs = (_ffi.HistoryField * max(1, len(fields)))()
    keep = []
    for i, (name, long_name, units, t, nlev) in enumerate(fields):
        _dev_field(name, t, ctx.ntiles * nlev * _slab(ctx))
        fs[i] = _ffi.HistoryField(name.encode(), long_name.encode(), units.encode(), t.data_ptr(), nlev,
                                  t.element_size())
    pf = np.ascontiguousarray(pfull, dtype=np.float32)
    ph = np.ascontiguousarray(phalf, dtype=np.float32)
    if pf.size != ctx.npz or ph.size != ctx.npz + 1:
        raise ValueError("pfull needs npz values and phalf npz + 1")

    def dp(a, n):
        if a is None:
            return None
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.size != n:
            raise ValueError(f"expected {n} values, got {a.size}")
        keep.append(a)
        return a.ctypes.data_as(_ffi.DP)

    b = ctx.bd
    ncell = ctx.ntiles * (b.ie - b.is_ + 1) * (b.je - b.js + 1)
    args = _ffi.HistoryArgs(len(fields), fs, out_bytes, quantize_nsd, tile_base,
                            pf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                            ph.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                            dp(lon, ncell), dp(lat, ncell), dp(ak, ctx.npz + 1), dp(bk, ctx.npz + 1),
                            time, time_units.encode(), time_iso.encode())
    check(_ffi.lib().fv3_history_write(ctx.handle, path.encode(), ctypes.byref(args), _stream(stream)),
          "fv3_history_write")


def lagrangian_to_eulerian_dev(ctx: Context, form, km, nq, ak, bk, ptop, state, kord_tr=9, kord_tm=10,
                               kord_mt=9, kord_wz=9, hydrostatic=False, consv=0.0, t_min=184.0,
                               fill=False, akap=287.05 / 1004.6, cp=1004.6, rdgas=287.05, grav=9.80665,
                               p_ref=1.0e5, mdt=0.0, last_step=False, fast_physics=None, stream=None):
    """fv_mapz.F90 Lagrangian_to_Eulerian on device tensors.  ``state`` maps
    delp, u, v, pt, pe, peln, pk, pkz, ps (and q, w, delz, ws, te0_2d when
    used) to contiguous device tensors in the layouts of
    fv3_lagrangian_to_eulerian_dev; field tensors are float64 for
    FORM_PLAIN64 and float32 for FORM_FUSED32, pressures and energies
    float64.  fast_physics (a Python callable taking the fv3_l2e_state) runs
    after the pressures and before the energy fix-up.  Returns dtmp."""
    import torch
    f = torch.float64 if form == FORM_PLAIN64 else torch.float32
    dbl = ("delp", "pe", "peln", "pk", "pkz", "ps", "te0_2d")
    for k, t in state.items():
        want = torch.float64 if k in dbl else f
        if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous() and t.dtype == want):
            raise TypeError(f"{k}: a contiguous {want} tensor on the HIP device is required")
    kt = np.full(max(nq, 1), kord_tr, dtype=np.int32)
    dtmp = ctypes.c_double(0.0)
    hook = _ffi.FAST_PHYSICS_FN(lambda user, st: int(bool(fast_physics(st.contents)))) if fast_physics \
        else _ffi.FAST_PHYSICS_FN()
    g = lambda k: ptr(state.get(k))
    a = _ffi.L2EArgs(km=km, nq=nq, ak=ptr(ak), bk=ptr(bk), ptop=ptop, akap=akap, cp=cp, rdgas=rdgas,
                     grav=grav, consv=consv, t_min=t_min, p_ref=p_ref, mdt=mdt,
                     hydrostatic=int(hydrostatic), fill=int(fill), kord_mt=kord_mt, kord_wz=kord_wz,
                     kord_tm=kord_tm, last_step=int(last_step), kord_tr=ptr(kt), delp=g("delp"),
                     u=g("u"), v=g("v"), w=g("w"), delz=g("delz"), pt=g("pt"), q=g("q"), ws=g("ws"),
                     pe=g("pe"), peln=g("peln"), pk=g("pk"), pkz=g("pkz"), ps=g("ps"),
                     te0_2d=g("te0_2d"), dtmp_out=ctypes.cast(ctypes.pointer(dtmp), ctypes.c_void_p),
                     fast_physics=hook, fast_physics_user=None)
    check(_ffi.lib().fv3_lagrangian_to_eulerian_dev(ctx.handle, form, ctypes.byref(a), _stream(stream)),
          "Lagrangian_to_Eulerian")
    return dtmp.value


def set_graph_safe(ctx: Context, on: bool = True):
    """Make fv3_step_dev capturable in a hipGraph (nsplit_cap > 0 required)."""
    check(_ffi.lib().fv3_ctx_set_graph_safe(ctx.handle, int(on)), "fv3_ctx_set_graph_safe")


def step_status(ctx: Context) -> int:
    """1 when a level of the last graph-safe step needed more sub-cycles than
    nsplit_cap (its results are not valid), else 0."""
    v = ctypes.c_int()
    check(_ffi.lib().fv3_step_status(ctx.handle, ctypes.byref(v)), "fv3_step_status")
    return v.value


def set_timing(ctx: Context, on: bool):
    """Record device times of the transport passes and the remap of each
    fused step (hipEvents on the step's stream)."""
    check(_ffi.lib().fv3_ctx_set_timing(ctx.handle, int(on)), "fv3_ctx_set_timing")


def step_timing(ctx: Context):
    """(transport_ms, remap_ms, passes) of the last timed fused step."""
    a, r, p = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
    check(_ffi.lib().fv3_step_timing(ctx.handle, ctypes.byref(a), ctypes.byref(r), ctypes.byref(p)),
          "fv3_step_timing")
    return a.value, r.value, p.value


def set_transport(ctx: Context, per_level: bool = True, q_split: int = 0):
    """Sub-cycling rule of step_dev: per-level nsplit (tracer_2d_1L, the
    default) or one nsplit for every level (tracer_2d), q_split if > 0."""
    check(_ffi.lib().fv3_ctx_set_transport(ctx.handle, 1 if per_level else 0, q_split),
          "fv3_ctx_set_transport")


def set_concurrency(ctx: Context, on: bool = True):
    """Edge kernels of the fused transport beside the lean kernel on auxiliary
    streams (True, the default) or after it on the step's stream."""
    check(_ffi.lib().fv3_ctx_set_concurrency(ctx.handle, 1 if on else 0), "fv3_ctx_set_concurrency")


def set_pipeline(ctx: Context, tracers_per_chunk: int = 0):
    """FP32 step: transport the tracers in chunks (-1: one round of the launch
    shape; 0, the default: all at once) and remap each chunk beside the next
    chunk's transport.  Results do not depend on the chunking."""
    check(_ffi.lib().fv3_ctx_set_pipeline(ctx.handle, int(tracers_per_chunk)), "fv3_ctx_set_pipeline")


DEFAULT_LAUNCH_CFG = 12


def set_launch(ctx: Context, cfg: int = DEFAULT_LAUNCH_CFG, rows: int = 0, remap: int = 0):
    """Kernel choices of the FP32 step: transport launch shape (cfg, the
    library default 12; rows) -- results unchanged -- and remap kernel (0
    column-shared with a column's levels across the lanes of a wave, 1 per
    column, 2 column-shared with one thread per (column, field))."""
    check(_ffi.lib().fv3_ctx_set_launch(ctx.handle, cfg, rows, remap), "fv3_ctx_set_launch")


N_LAUNCH_CFG = 15
