"""The on-disk formats' per-value arithmetic, restated in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker; the product
never imports it.

- mpp_chksum: FMS's restart checksum of a field's compute domain, the sum
  modulo 2**64 of the values' bit patterns read as two's-complement
  integers (FMS mpp_chksum, called by fms2_io when it writes the "checksum"
  attribute of each restart variable; FMS is not part of the reference, so
  the restatement follows its documented behaviour).  The reference reads
  the attribute back unless ignore_rst_cksum (atmos_model.F90:193-194).
- cksum_text: the attribute text, a Fortran Z16 edit (upper-case hex,
  right-justified in 16 characters).
- bitround: netCDF's quantize_bitround (io/module_write_netcdf.F90:417-443,
  nsd = mantissa bits kept): the dropped mantissa bits are rounded to
  nearest by adding half of the last kept bit to the magnitude and
  truncating; NaN and infinities pass through.  netCDF-C is not part of the
  reference: how a tie rounds (away from zero here) is parity unpinned.
- interior: the compute domain (is:ie, js:je) of a halo'd (isd:ied, jsd:jed,
  ...) array.
"""
from __future__ import annotations

import numpy as np

NG = 3


def mpp_chksum(a) -> int:
    a = np.ascontiguousarray(a)
    if a.dtype == np.float64:
        bits = a.view(np.int64)
    elif a.dtype == np.float32:
        bits = a.view(np.int32).astype(np.int64)
    else:
        raise TypeError(a.dtype)
    # int64 sums wrap modulo 2**64 in numpy; the result as unsigned
    with np.errstate(over="ignore"):
        s = np.sum(bits, dtype=np.int64)
    return int(s) & (2**64 - 1)


def cksum_text(v: int) -> str:
    return f"{v:16X}"


def bitround(a, nsd: int):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        u = a.view(np.uint32).copy()
        mb, exp = 23, np.uint32(0x7F800000)
        one = np.uint32(1)
    elif a.dtype == np.float64:
        u = a.view(np.uint64).copy()
        mb, exp = 52, np.uint64(0x7FF0000000000000)
        one = np.uint64(1)
    else:
        raise TypeError(a.dtype)
    if nsd <= 0 or nsd >= mb:
        return a.copy()
    sh = mb - nsd
    half = one << type(one)(sh - 1)
    mask = ~((one << type(one)(sh)) - one)
    fin = (u & exp) != exp
    u[fin] = (u[fin] + half) & mask
    return u.view(a.dtype)


def interior(a, ng=NG):
    return a[ng:-ng, ng:-ng]
