"""ctypes binding of the C-ABI declared in include/fv3atm_amd.h.

The product path has no CPU fallback: if libfv3atm_amd.so is missing or a
call fails, a RuntimeError is raised with the library's own error text.
"""
from __future__ import annotations

import ctypes
import os

_here = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_here, "libfv3atm_amd.so")

P = ctypes.c_void_p
I = ctypes.c_int
D = ctypes.c_double
IP = ctypes.POINTER(ctypes.c_int)
DP = ctypes.POINTER(ctypes.c_double)

FORM_PLAIN64, FORM_FUSED64, FORM_FUSED32 = 0, 1, 2
NOALLOC_SCRATCH_BYTES = 256 << 20  # FV3_NOALLOC_SCRATCH_BYTES
AGRID, CGRID, DGRID = 1, 2, 3


class Bounds(ctypes.Structure):
    _fields_ = [(k, I) for k in ("is_", "ie", "js", "je", "isd", "ied", "jsd", "jed")]


class Grid(ctypes.Structure):
    _fields_ = [(k, P) for k in ("area", "rarea", "dx", "dy", "dxa", "dya", "sin_sg",
                                 "del6_u", "del6_v")] + [("da_min", D)] + \
               [(k, I) for k in ("sw_corner", "se_corner", "nw_corner", "ne_corner",
                                 "bounded_domain", "grid_type")]


class L2EState(ctypes.Structure):
    _fields_ = [(k, I) for k in ("npz", "nq", "ntiles", "kmp", "last_step")] + [("mdt", D)] + \
               [(k, P) for k in ("delp", "pe", "peln", "pk", "pkz", "ps", "te0_2d", "pt", "q", "delz",
                                 "w", "stream")]


FAST_PHYSICS_FN = ctypes.CFUNCTYPE(ctypes.c_int, P, ctypes.POINTER(L2EState))


class HistoryField(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char_p), ("long_name", ctypes.c_char_p), ("units", ctypes.c_char_p),
                ("data", P), ("nlev", I), ("elem_bytes", I)]


class HistoryArgs(ctypes.Structure):
    _fields_ = [("nfields", I), ("fields", ctypes.POINTER(HistoryField)), ("out_bytes", I),
                ("quantize_nsd", I), ("tile_base", I), ("pfull", ctypes.POINTER(ctypes.c_float)),
                ("phalf", ctypes.POINTER(ctypes.c_float))] + \
               [(k, DP) for k in ("lon", "lat", "ak", "bk")] + \
               [("time", D), ("time_units", ctypes.c_char_p), ("time_iso", ctypes.c_char_p)]


class L2EArgs(ctypes.Structure):
    _fields_ = [("km", I), ("nq", I), ("ak", P), ("bk", P)] + \
               [(k, D) for k in ("ptop", "akap", "cp", "rdgas", "grav", "consv", "t_min", "p_ref", "mdt")] + \
               [(k, I) for k in ("hydrostatic", "fill", "kord_mt", "kord_wz", "kord_tm", "last_step")] + \
               [(k, P) for k in ("kord_tr", "delp", "u", "v", "w", "delz", "pt", "q", "ws", "pe", "peln",
                                 "pk", "pkz", "ps", "te0_2d", "dtmp_out")] + \
               [("fast_physics", FAST_PHYSICS_FN), ("fast_physics_user", P)]


# every symbol the header declares, with its ctypes signature
SIGNATURES = {
    "fv3_last_error": (ctypes.c_char_p, []),
    "fv3_ctx_create": (I, [ctypes.POINTER(Bounds), I, I, I, I, ctypes.POINTER(P)]),
    "fv3_ctx_destroy": (I, [P]),
    "fv3_ctx_set_grid": (I, [P, I, ctypes.POINTER(Grid)]),
    "fv3_fv_tp_2d": (I, [P, P, P, P, I, I, I, P, P, P, P, P, P, D, P, P, P, IP, DP]),
    "fv3_tracer_2d_1l": (I, [P, P, P, P, P, P, P, I, I, I, I, I, I, D, I, I, D, D, P]),
    "fv3_tracer_2d": (I, [P, P, P, P, P, P, P, I, I, I, I, I, I, D, I, I, D, D, P]),
    "fv3_tracer_2d_1l_r4": (I, [P, P, P, P, P, P, P, I, I, I, I, I, I, D, I, I, D, D, P]),
    "fv3_tracer_2d_r4": (I, [P, P, P, P, P, P, P, I, I, I, I, I, I, D, I, I, D, D, P]),
    "fv3_lagrangian_to_eulerian_tracers_r4": (I, [P, I, I, P, P, D, P, P, P, P, I, D, I]),
    "fv3_map1_ppm": (I, [I, P, P, P, I, P, P, I, I, I, I, I, I, I, I, I]),
    "fv3_map_scalar": (I, [I, P, P, P, I, P, P, I, I, I, I, I, I, I, I, I, D]),
    "fv3_mapn_tracer": (I, [I, I, P, P, P, P, P, I, I, I, I, I, I, I, D, I]),
    "fv3_cs_profile": (I, [P, P, P, I, I, I, I, I]),
    "fv3_lagrangian_to_eulerian_tracers": (I, [P, I, I, P, P, D, P, P, P, P, I, D, I]),
    "fv3_step_dev": (I, [P, I, I, P, P, P, P, P, P, P, P, P, P, P, D, I, I, I, D, I, P]),
    "fv3_step_dev_noalloc": (I, [P, I, I, P, P, P, P, P, P, P, P, P, P, P, D, I, I, I, D, I, P]),
    "fv3_ctx_set_timing": (I, [P, I]),
    "fv3_step_timing": (I, [P, DP, DP, IP]),
    "fv3_ctx_set_launch": (I, [P, I, I, I]),
    "fv3_ctx_set_concurrency": (I, [P, I]),
    "fv3_ctx_set_pipeline": (I, [P, I]),
    "fv3_ctx_set_transport": (I, [P, I, I]),
    "fv3_ctx_set_graph_safe": (I, [P, I]),
    "fv3_step_status": (I, [P, IP]),
    "fv3_halo_update": (I, [P, P, I, I, P]),
    "fv3_halo_update_vector": (I, [P, P, P, I, I, I, P]),
    "fv3_lagrangian_to_eulerian_dev": (I, [P, I, ctypes.POINTER(L2EArgs), P]),
    "fv3_kmp": (I, [I, D, P, P, IP]),
    "fv3_dyn_core_fluxes": (I, [P, I, I, D, I, P, P, P, P, P, P, P, P, P]),
    "fv3_global_mass": (I, [P, P, I, P, I, DP, P]),
    "fv3_field_checksum": (I, [P, P, I, I, ctypes.POINTER(ctypes.c_ulonglong), P]),
    "fv3_tracer_restart_write": (I, [P, ctypes.c_char_p, I, I, ctypes.POINTER(ctypes.c_char_p), P, I, P]),
    "fv3_tracer_restart_read": (I, [P, ctypes.c_char_p, I, I, ctypes.POINTER(ctypes.c_char_p), P, I, I, P]),
    "fv3_history_write": (I, [P, ctypes.c_char_p, ctypes.POINTER(HistoryArgs), P]),
    "fv3_step_tl_dev": (I, [P, I, I, I, P, P, P, P, P, P, P, P, P, P, P, P, D, I, I, P]),
    "fv3_step_ad_dev": (I, [P, I, I, I, P, P, P, P, P, P, P, P, P, P, D, I, I, P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().fv3_last_error()
        raise RuntimeError(f"{what} failed: {msg.decode() if msg else rc}")


def ptr(a):
    """Address of a numpy array (host) or torch tensor (device); None -> NULL."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return ctypes.c_void_p(a.data_ptr())
    return a.ctypes.data_as(ctypes.c_void_p)
