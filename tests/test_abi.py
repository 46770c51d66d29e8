"""The C-ABI library loads and exports every symbol include/fv3atm_amd.h
declares, with no GPU call; argument validation that needs no device runs
on the host and returns a nonzero status with a reason."""
import ctypes
import os
import re

import numpy as np
import pytest

from fv3atm_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "fv3atm_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fv3_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    syms = header_symbols()
    assert len(syms) >= 14
    lib = ctypes.CDLL(_ffi.LIB_PATH)
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding covers the same set
    assert set(syms) == set(_ffi.SIGNATURES), set(syms) ^ set(_ffi.SIGNATURES)


def test_row_entry_validation_on_host():
    lib = _ffi.lib()
    a = np.zeros(64)
    # kord 7 (ppm_profile) is not part of this path: refused before any device work
    rc = lib.fv3_map1_ppm(8, _ffi.ptr(a), _ffi.ptr(a), _ffi.ptr(a), 8, _ffi.ptr(a), _ffi.ptr(a),
                          1, 2, 1, 1, 2, 1, 1, 0, 7)
    assert rc != 0 and b"kord" in lib.fv3_last_error()
    rc = lib.fv3_map1_ppm(8, None, _ffi.ptr(a), None, 8, _ffi.ptr(a), _ffi.ptr(a), 1, 2, 1, 1, 2,
                          1, 1, 0, 9)
    assert rc != 0 and b"NULL" in lib.fv3_last_error()
    rc = lib.fv3_mapn_tracer(0, 8, _ffi.ptr(a), _ffi.ptr(a), _ffi.ptr(a), _ffi.ptr(a),
                             _ffi.ptr(np.zeros(1, np.int32)), 1, 1, 2, 1, 2, 1, 1, 0.0, 0)
    assert rc != 0 and b"nq" in lib.fv3_last_error()


def test_context_validation_on_host():
    lib = _ffi.lib()
    h = ctypes.c_void_p()
    bd = _ffi.Bounds(1, 16, 1, 16, -1, 18, -1, 18)  # 2-cell halo
    assert lib.fv3_ctx_create(ctypes.byref(bd), 17, 17, 4, 1, ctypes.byref(h)) != 0
    assert b"halo" in lib.fv3_last_error()
    bd = _ffi.Bounds(1, 16, 1, 16, -2, 19, -2, 19)
    assert lib.fv3_ctx_create(ctypes.byref(bd), 17, 17, 4, 3, ctypes.byref(h)) != 0
    assert b"ntiles" in lib.fv3_last_error()
    assert lib.fv3_step_dev(None, 0, 1, None, None, None, None, None, None, None, None, None,
                            None, None, 0.0, 8, 9, 10, 184.0, 0, None) != 0
