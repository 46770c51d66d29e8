"""Pins of the on-disk formats' oracle (oracle/io_oracle.py): known answers
and what the arithmetic fixes."""
import numpy as np

from oracle import io_oracle as io


def test_checksum_known_answers():
    # 1.0 = 0x3FF0..., 2.0 = 0x4000...: their sum is the bit pattern of +inf
    assert io.mpp_chksum(np.array([1.0, 2.0])) == 0x7FF0000000000000
    # -0.0 is the most negative int64: two of them wrap to zero
    assert io.mpp_chksum(np.array([-0.0, -0.0])) == 0
    assert io.mpp_chksum(np.array([-0.0])) == 0x8000000000000000
    # float32 patterns are sign-extended: -1.0f = 0xBF800000 -> 0xFFFFFFFFBF800000
    assert io.mpp_chksum(np.array([-1.0], dtype=np.float32)) == 0xFFFFFFFFBF800000
    assert io.mpp_chksum(np.array([1.0, -1.0], dtype=np.float32)) == (0x3F800000 + 0xFFFFFFFFBF800000) % 2**64
    assert io.cksum_text(0x7FF0000000000000) == "7FF0000000000000"
    assert io.cksum_text(0xABC) == "             ABC"


def test_checksum_is_order_free_and_additive():
    rng = np.random.default_rng(1)
    a = rng.standard_normal(1000) * 1e5
    b = rng.standard_normal(500)
    ca, cb = io.mpp_chksum(a), io.mpp_chksum(b)
    assert io.mpp_chksum(np.concatenate([a, b])) == (ca + cb) % 2**64
    assert io.mpp_chksum(rng.permutation(a)) == ca
    # a one-ulp change moves it by exactly one
    a2 = a.copy()
    a2[17] = np.nextafter(a2[17], np.inf)
    assert (io.mpp_chksum(a2) - ca) % 2**64 == (1 if a[17] > 0 else 2**64 - 1)


def test_bitround_known_answers():
    f = np.float32
    x = np.array([1.0, 1.5, 1 + 2**-23, 1 + 2**-11, 1 + 2**-11 - 2**-23, -1 - 2**-11, np.inf, -np.inf], dtype=f)
    r = io.bitround(x, 10)
    # nsd = 10 keeps mantissa bits down to 2**-10; 2**-11 is the half
    np.testing.assert_array_equal(r[:6], np.array([1.0, 1.5, 1.0, 1 + 2**-10, 1.0, -1 - 2**-10], dtype=f))
    assert np.isinf(r[6]) and r[6] > 0 and np.isinf(r[7]) and r[7] < 0
    assert np.isnan(io.bitround(np.array([np.nan], dtype=f), 10)[0])
    # nsd 0 or >= the mantissa width leaves values alone
    np.testing.assert_array_equal(io.bitround(x[:6], 0), x[:6])
    np.testing.assert_array_equal(io.bitround(x[:6], 23), x[:6])
    # a carry out of the mantissa moves to the next binade
    y = np.array([2.0 - 2**-20], dtype=f)
    assert io.bitround(y, 8)[0] == 2.0


def test_bitround_properties():
    rng = np.random.default_rng(2)
    for dt, mb in ((np.float32, 23), (np.float64, 52)):
        x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000)).astype(dt)
        for nsd in (3, 7, 12, 20):
            r = io.bitround(x, nsd)
            ubits = r.view(np.uint32 if dt == np.float32 else np.uint64)
            low = (ubits & ((1 << (mb - nsd)) - 1)) if dt == np.float32 else (ubits & np.uint64((1 << (mb - nsd)) - 1))
            assert not low.any()                                   # only nsd mantissa bits left
            rel = np.abs(r.astype(np.float64) - x.astype(np.float64)) / np.abs(x.astype(np.float64))
            assert rel.max() <= 2.0 ** -(nsd + 1) * (1 + 1e-12)     # half a unit of the last kept bit
            np.testing.assert_array_equal(io.bitround(r, nsd), r)   # idempotent
            assert (np.sign(r) == np.sign(x)).all()
            o = np.argsort(x)
            assert (np.diff(r[o].astype(np.float64)) >= 0).all()   # monotone
