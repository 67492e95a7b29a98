"""Pins the oracle's radix-sort semantics (tinyhiponesweep.h:592-668) against
numpy's stable argsort on the masked digit field."""
import numpy as np
import pytest


def masked_field(keys, start, end, bits):
    passes = (end - start + 7) // 8
    hi = min(start + 8 * passes, bits)
    k = keys.astype(np.uint64)
    width = hi - start
    if width <= 0:
        return np.zeros_like(k)
    mask = np.uint64((1 << width) - 1) if width < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    return (k >> np.uint64(start)) & mask


@pytest.mark.parametrize("dtype,bits", [(np.uint32, 32), (np.uint64, 64)])
@pytest.mark.parametrize("rng_", [(0, None), (0, 30), (8, 40), (4, 12), (13, 14), (24, 32)])
def test_radix_matches_stable_argsort(orc, dtype, bits, rng_):
    start, end = rng_
    end = bits if end is None else end
    r = np.random.default_rng(start * 100 + end)
    n = 5003
    keys = r.integers(0, 2**bits if bits < 64 else 2**63, size=n, dtype=np.uint64).astype(dtype)
    keys[::7] = keys[3]  # ties
    vals = np.arange(n, dtype=np.uint32)
    sk, sv, passes = orc.radix_sort(keys, vals, start, end)
    assert passes == (end - start + 7) // 8
    order = np.argsort(masked_field(keys, start, end, bits), kind="stable")
    assert np.array_equal(sv, vals[order])
    assert np.array_equal(sk, keys[order])


def test_radix_edge_sizes(orc):
    for n in [0, 1, 2, 255, 256, 257]:
        keys = np.random.default_rng(n).integers(0, 2**64 - 1, size=n, dtype=np.uint64)
        sk, _, _ = orc.radix_sort(keys)
        assert np.array_equal(sk, np.sort(keys))


def test_radix_rejects_bad_range(orc):
    _, _, p = orc.radix_sort(np.zeros(4, np.uint32), None, 8, 8)
    assert p == -1
