"""The on-disk formats (fv3_tracer_restart_write/read, fv3_history_write,
fv3_field_checksum) on the GPU.  Files are read with scipy.io.netcdf_file,
an independent implementation of the netCDF classic format; values are
compared bit for bit with the device arrays' compute domains, checksums and
BitRound with the numpy restatement (oracle/io_oracle.py), and a restart
written by scipy is read back through the library."""
import os

import numpy as np
import pytest
import scipy.io
import torch

import fv3atm_amd as fa
from fv3atm_amd import cases
from oracle import io_oracle as io

pytestmark = pytest.mark.gpu
NG = 3
NAMES = ["sphum", "liq_wat", "o3mr"]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).cuda()


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def _q(n, npz, nq, nt, dtype, seed):
    rng = np.random.default_rng(seed)
    nd = n + 2 * NG
    q = rng.standard_normal((nd, nd, npz, nq, nt)) * np.logspace(-6, 2, nq)[None, None, None, :, None]
    return np.asfortranarray(q.astype(dtype))


def _ctx(n, npz, ntiles):
    case = cases.make_case(f"C{n}L{npz}", n=n, npz=npz, nq=1, ntiles=ntiles, seed=1)
    return fa.Context(case.grids, npz), case


@pytest.mark.parametrize("ntiles,dtype", [(1, np.float64), (6, np.float32)])
def test_restart_round_trip(gpu_ready, tmp_path, ntiles, dtype):
    n, npz = 12, 5
    print(f"[io] restart {ntiles} tile(s) {np.dtype(dtype).name}", flush=True)
    ctx, _ = _ctx(n, npz, ntiles)
    q = _q(n, npz, len(NAMES), ntiles, dtype, 3)
    dq = _dev(q)
    prefix = str(tmp_path / "fv_tracer.res")
    fa.tracer_restart_write(ctx, prefix, NAMES, dq)
    for t in range(ntiles):
        f = scipy.io.netcdf_file(f"{prefix}.tile{t + 1}.nc", "r", mmap=False)
        assert f.version_byte == 2
        assert f.dimensions == {"xaxis_1": n, "yaxis_1": n, "zaxis_1": npz, "Time": None}
        np.testing.assert_array_equal(f.variables["xaxis_1"][:], np.arange(1, n + 1))
        np.testing.assert_array_equal(f.variables["zaxis_1"][:], np.arange(1, npz + 1))
        np.testing.assert_array_equal(f.variables["Time"][:], [1.0])
        assert f.variables["Time"].cartesian_axis == b"T"
        for iq, name in enumerate(NAMES):
            v = f.variables[name]
            assert v.dimensions == ("Time", "zaxis_1", "yaxis_1", "xaxis_1")
            assert v.typecode() == "d" and v.long_name == name.encode() and v.units == b"none"
            want = io.interior(q[..., iq, t]).astype(np.float64)        # (i, j, k)
            got = np.transpose(v[0], (2, 1, 0))
            np.testing.assert_array_equal(got, want)
            assert v.checksum.decode() == io.cksum_text(io.mpp_chksum(want.ravel(order="F")))
        f.close()
    # read back into an array whose halo holds a sentinel: the compute domain
    # comes back exactly, the halo is untouched
    back = torch.full_like(dq, -7.0)
    fa.tracer_restart_read(ctx, prefix, NAMES, back)
    b = _host(back, q.shape)
    np.testing.assert_array_equal(io.interior(b), io.interior(q))
    halo = np.ones(b.shape[:2], bool)
    halo[NG:-NG, NG:-NG] = False
    assert (b[halo] == -7.0).all()
    # a subset of the tracers, in another order
    sub = torch.zeros(ntiles * npz * (n + 6) ** 2 * 2, dtype=dq.dtype, device="cuda")
    fa.tracer_restart_read(ctx, prefix, ["o3mr", "sphum"], sub)
    s = _host(sub, q.shape[:3] + (2, ntiles))
    np.testing.assert_array_equal(io.interior(s[..., 0, :]), io.interior(q[..., 2, :]))
    np.testing.assert_array_equal(io.interior(s[..., 1, :]), io.interior(q[..., 0, :]))
    ctx.close()


def test_restart_checksum_guards_the_data(gpu_ready, tmp_path):
    n, npz = 10, 4
    ctx, _ = _ctx(n, npz, 1)
    q = _q(n, npz, 2, 1, np.float64, 5)
    prefix = str(tmp_path / "fv_tracer.res")
    fa.tracer_restart_write(ctx, prefix, NAMES[:2], _dev(q))
    path = prefix + ".tile1.nc"
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    # locate one value of liq_wat in the file and change its last bit
    raw = bytearray(open(path, "rb").read())
    needle = np.array([io.interior(q[..., 1, 0])[2, 3, 1]], dtype=">f8").tobytes()
    at = raw.find(needle)
    assert at > 0
    f.close()
    raw[at + 7] ^= 1
    open(path, "wb").write(bytes(raw))
    back = torch.zeros(npz * (n + 6) ** 2 * 2, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="checksum of liq_wat"):
        fa.tracer_restart_read(ctx, prefix, NAMES[:2], back)
    fa.tracer_restart_read(ctx, prefix, NAMES[:2], back, ignore_cksum=True)
    b = _host(back, q.shape[:4])
    assert (io.interior(b[..., 1]) != io.interior(q[..., 1, 0])).sum() == 1
    with pytest.raises(RuntimeError, match="no variable"):
        fa.tracer_restart_read(ctx, prefix, ["ice_wat"], back)
    with pytest.raises(RuntimeError, match="cannot open"):
        fa.tracer_restart_read(ctx, str(tmp_path / "absent"), NAMES[:1], back)
    ctx.close()


def test_restart_written_by_scipy(gpu_ready, tmp_path):
    # a CDF-1 file from an independent writer, with the FMS checksum computed
    # by the oracle, goes through the library's reader
    n, npz = 8, 3
    ctx, _ = _ctx(n, npz, 1)
    rng = np.random.default_rng(9)
    data = rng.standard_normal((npz, n, n))                   # (k, j, i)
    prefix = str(tmp_path / "ext")
    f = scipy.io.netcdf_file(prefix + ".tile1.nc", "w", version=1)
    f.createDimension("Time", None)
    for d, m in (("xaxis_1", n), ("yaxis_1", n), ("zaxis_1", npz)):
        f.createDimension(d, m)
    v = f.createVariable("sphum", "d", ("Time", "zaxis_1", "yaxis_1", "xaxis_1"))
    v[0] = data
    v.checksum = io.cksum_text(io.mpp_chksum(np.transpose(data, (2, 1, 0)).ravel(order="F")))
    f.close()
    back = torch.zeros(npz * (n + 6) ** 2, dtype=torch.float32, device="cuda")
    fa.tracer_restart_read(ctx, prefix, ["sphum"], back)
    b = _host(back, (n + 6, n + 6, npz))
    np.testing.assert_array_equal(io.interior(b), np.transpose(data, (2, 1, 0)).astype(np.float32))
    ctx.close()


@pytest.mark.parametrize("nt", [1, 6])
def test_field_checksum(gpu_ready, nt):
    n, npz = 14, 3
    ctx, _ = _ctx(n, npz, nt)
    for dtype in (np.float64, np.float32):
        q = _q(n, npz, 1, nt, dtype, 11)[..., 0, :]
        got = fa.field_checksum(ctx, _dev(q), npz)
        want = [io.mpp_chksum(io.interior(q[..., t]).ravel(order="F")) for t in range(nt)]
        assert got == want, dtype
    ctx.close()


@pytest.mark.parametrize("out_bytes,nsd", [(4, 10), (8, 0), (4, 0)])
def test_history_bundle(gpu_ready, tmp_path, out_bytes, nsd):
    n, npz, nt = 12, 6, 6
    print(f"[io] history out_bytes={out_bytes} nsd={nsd}", flush=True)
    ctx, case = _ctx(n, npz, nt)
    rng = np.random.default_rng(4)
    nd = n + 2 * NG
    q = np.asfortranarray((rng.random((nd, nd, npz, nt)) * 1e-3).astype(np.float32))
    pt = np.asfortranarray(200.0 + 100.0 * rng.random((nd, nd, npz, nt)))
    ps = np.asfortranarray(1e5 + 100.0 * rng.standard_normal((nd, nd, 1, nt)))
    lon = rng.random((nt, n, n)) * 360.0
    lat = rng.random((nt, n, n)) * 180.0 - 90.0
    pfull = np.linspace(10.0, 1000.0, npz)
    phalf = np.linspace(5.0, 1005.0, npz + 1)
    path = str(tmp_path / "dynf000.nc")
    fa.history_write(ctx, path, [("spfh", "specific humidity", "kg/kg", _dev(q), npz),
                                 ("tmp", "temperature", "K", _dev(pt), npz),
                                 ("pressfc", "surface pressure", "pa", _dev(ps), 1)],
                     pfull, phalf, out_bytes=out_bytes, quantize_nsd=nsd, lon=lon, lat=lat,
                     ak=case.ak, bk=case.bk, time=3.0, time_iso="2026-12-03T03:00:00Z")
    f = scipy.io.netcdf_file(path, "r", mmap=False)
    assert f.version_byte == 2
    assert f.dimensions == {"grid_xt": n, "grid_yt": n, "nchars": 20, "pfull": npz, "phalf": npz + 1,
                            "tile": nt, "time": 1}
    np.testing.assert_array_equal(f.variables["grid_xt"][:], np.arange(1, n + 1))
    np.testing.assert_array_equal(f.variables["tile"][:], np.arange(1, nt + 1))
    np.testing.assert_array_equal(f.variables["pfull"][:], pfull.astype(np.float32))
    np.testing.assert_array_equal(f.variables["phalf"][:], phalf.astype(np.float32))
    np.testing.assert_array_equal(f.variables["time"][:], [3.0])
    assert b"".join(f.variables["time_iso"][0]).rstrip(b"\0") == b"2026-12-03T03:00:00Z"
    np.testing.assert_array_equal(f.variables["lon"][:], lon)
    np.testing.assert_array_equal(f.variables["lat"][:], lat)
    np.testing.assert_array_equal(f.ak, case.ak)
    assert f.grid == b"cubed_sphere" and f.im == n and f.jm == n
    ftype = np.float32 if out_bytes == 4 else np.float64
    for name, src, nlev in (("spfh", q, npz), ("tmp", pt, npz), ("pressfc", ps[..., 0, :], 1)):
        v = f.variables[name]
        dims = ("time", "tile", "pfull", "grid_yt", "grid_xt") if nlev > 1 else ("time", "tile", "grid_yt", "grid_xt")
        assert v.dimensions == dims
        assert v.coordinates == b"lon lat" and v.grid_mapping == b"cubed_sphere"
        val = io.interior(src).astype(ftype)
        if nlev > 1 and nsd > 0:
            val = io.bitround(val, nsd)
            assert v._QuantizeBitRoundNumberOfSignificantBits == nsd
        else:
            assert not hasattr(v, "_QuantizeBitRoundNumberOfSignificantBits")
        want = np.transpose(val, (3, 2, 1, 0)) if nlev > 1 else np.transpose(val, (2, 1, 0))
        got = v[0]
        assert got.dtype == np.dtype(ftype).newbyteorder(">") or got.dtype == ftype
        np.testing.assert_array_equal(got.astype(ftype), want)
    f.close()
    ctx.close()


def test_io_argument_checks(gpu_ready, tmp_path):
    ctx, _ = _ctx(8, 3, 1)
    q = torch.zeros(3 * 14 * 14, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="quantize_nsd"):
        fa.history_write(ctx, str(tmp_path / "h.nc"), [("a", "a", "1", q, 3)], np.ones(3), np.ones(4),
                         quantize_nsd=30)
    with pytest.raises(RuntimeError, match="nlev"):
        fa.history_write(ctx, str(tmp_path / "h.nc"), [("a", "a", "1", q, 2)], np.ones(3), np.ones(4))
    with pytest.raises(RuntimeError, match="cannot create"):
        fa.tracer_restart_write(ctx, str(tmp_path / "no" / "dir" / "x"), ["a"], q)
    ctx.close()
