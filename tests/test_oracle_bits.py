"""Pins the oracle's bit utilities: the reference's own unit tests
(unittest.cpp:6-105) plus closed forms."""
import numpy as np


def test_clz_edge_cases(orc):  # unittest.cpp:39-74
    L = orc.lib()
    assert L.orc_clz(0) == 32
    for i in range(32):
        x = 0x80000000 >> i
        assert L.orc_clz(x) == i
        assert L.orc_clz(x | (x - 1)) == i
    assert L.orc_clz64(0) == 64
    for i in range(64):
        x = 0x8000000000000000 >> i
        assert L.orc_clz64(x) == i
        assert L.orc_clz64(x | (x - 1)) == i


def test_ffs_edge_cases(orc):  # unittest.cpp:76-105
    L = orc.lib()
    assert L.orc_ffs(0) == 0
    for i in range(32):
        x = 1 << i
        assert L.orc_ffs(x) == i + 1
        assert L.orc_ffs((x | ~(x - 1)) & 0xFFFFFFFF) == i + 1
    assert L.orc_ffs64(0) == 0
    for i in range(64):
        x = 1 << i
        assert L.orc_ffs64(x) == i + 1
        assert L.orc_ffs64((x | ~(x - 1)) & 0xFFFFFFFFFFFFFFFF) == i + 1


def test_delta_closed_form(orc):
    rng = np.random.default_rng(3)
    a = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    b = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    b[:100] = a[:100]  # equal keys -> delta 0
    for x, y in zip(a.tolist(), b.tolist()):
        expect = (x ^ y).bit_length()  # 32 - clz(x ^ y)
        assert orc.lib().orc_delta32(x, y) == expect


def test_splat3_exhaustive(orc):  # unittest.cpp:10-19: all 2^21 values
    x = np.arange(1 << 21, dtype=np.uint32)
    xyz = np.zeros((x.size, 3), np.uint32)
    xyz[:, 0] = x
    s = orc.morton_batch(xyz)
    rebuilt = np.zeros_like(s)
    for i in range(21):
        bit = (s >> np.uint64(3 * i)) & np.uint64(1)
        assert np.array_equal(bit, ((x >> i) & 1).astype(np.uint64))
        rebuilt |= bit << np.uint64(3 * i)
    assert np.array_equal(rebuilt, s)  # no stray bits
    assert np.array_equal(orc.decode_batch(s)[:, 0], x)


def test_morton_naive_equals_fast_and_roundtrip(orc):  # unittest.cpp:21-36
    rng = np.random.default_rng(11)
    xyz = rng.integers(0, 1 << 21, size=(2_000_000, 3), dtype=np.uint32)
    fast = orc.morton_batch(xyz)
    naive = orc.morton_batch(xyz, naive=True)
    assert np.array_equal(fast, naive)
    assert np.array_equal(orc.decode_batch(fast), xyz)
    # independent numpy interleave: bit i of x/y/z -> bit 3i/3i+1/3i+2
    ref = np.zeros(xyz.shape[0], np.uint64)
    for i in range(21):
        for c in range(3):
            ref |= ((xyz[:, c].astype(np.uint64) >> np.uint64(i)) & np.uint64(1)) << np.uint64(3 * i + c)
    assert np.array_equal(fast, ref)


def test_f2i_truncation_rule(orc):
    L = orc.lib()
    assert L.orc_f2i_trunc(2.9) == 2
    assert L.orc_f2i_trunc(-2.9) == -2
    assert L.orc_f2i_trunc(float("nan")) == -(2**31)
    assert L.orc_f2i_trunc(float("inf")) == -(2**31)
    assert L.orc_f2i_trunc(3e9) == -(2**31)
    assert L.orc_f2i_trunc(-2147483648.0) == -(2**31)
    assert L.orc_f2i_trunc(2097152.0) == 2097152
