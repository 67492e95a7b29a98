"""GPU Onesweep sort (mlbvh_sort_u32/u64) vs the oracle's stable LSD radix
sort: bit-exact keys and values."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _dev(a):
    a = np.ascontiguousarray(a)
    view = {np.dtype(np.uint64): np.int64, np.dtype(np.uint32): np.int32}.get(a.dtype, a.dtype)
    return torch.from_numpy(a.view(view)).cuda()


def _host(t, dtype):
    return t.cpu().numpy().view(dtype)


def _run(gpu, keys, vals, start, end):
    n = keys.shape[0]
    k = _dev(keys)
    kt = torch.empty_like(k)
    v = _dev(vals) if vals is not None else None
    vt = torch.empty_like(v) if v is not None else None
    gpu.OnesweepSort().sort(k, v, kt, vt, n, start, end)
    torch.cuda.synchronize()
    return _host(k, keys.dtype), (_host(v, np.uint32) if v is not None else None)


@pytest.mark.parametrize("n", [0, 1, 2, 100, 4095, 4096, 4097, 65537, (1 << 20) + 3])
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 32), (np.uint64, 64)])
def test_sort_full_range(gpu, orc, n, dtype, bits):
    print(f"sort n={n} bits={bits}", flush=True)
    r = np.random.default_rng(n + bits)
    keys = r.integers(0, np.iinfo(dtype).max, size=n, dtype=dtype, endpoint=True)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = _run(gpu, keys, vals, 0, bits)
    ok, ov, _ = orc.radix_sort(keys, vals, 0, bits)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)


@pytest.mark.parametrize("rng_", [(0, 30), (8, 40), (4, 12), (24, 32), (13, 14), (0, 8)])
@pytest.mark.parametrize("dtype,bits", [(np.uint32, 32), (np.uint64, 64)])
@pytest.mark.parametrize("with_values", [False, True])
def test_sort_bit_ranges(gpu, orc, rng_, dtype, bits, with_values):
    start, end = rng_
    print(f"sort range {start}-{end} bits={bits} values={with_values}", flush=True)
    n = 300_007
    r = np.random.default_rng(start * 1000 + end)
    keys = r.integers(0, np.iinfo(dtype).max, size=n, dtype=dtype, endpoint=True)
    keys[::5] = keys[1]  # ties must keep input order
    vals = r.integers(0, 2**32, size=n, dtype=np.uint32) if with_values else None
    gk, gv = _run(gpu, keys, vals, start, end)
    ok, ov, _ = orc.radix_sort(keys, vals, start, end)
    assert np.array_equal(gk, ok)
    if with_values:
        assert np.array_equal(gv, ov)


@pytest.mark.parametrize("pattern", ["equal", "sorted", "reversed", "two"])
def test_sort_patterns(gpu, orc, pattern):
    n = 200_000
    if pattern == "equal":
        keys = np.full(n, 0xDEADBEEFCAFE, np.uint64)
    elif pattern == "sorted":
        keys = np.arange(n, dtype=np.uint64) * 977
    elif pattern == "reversed":
        keys = (np.arange(n, dtype=np.uint64) * 977)[::-1].copy()
    else:
        keys = (np.arange(n, dtype=np.uint64) % 2) << np.uint64(40)
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = _run(gpu, keys, vals, 0, 64)
    ok, ov, _ = orc.radix_sort(keys, vals, 0, 64)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


def test_sort_repeated_calls_and_shrinking_sizes(gpu, orc):
    """Look-back words are tagged per pass; back-to-back sorts of different
    sizes on one context must not see each other's state."""
    for n in [500_000, 70_000, 4097, 1_000_003, 3]:
        keys = np.random.default_rng(n).integers(0, 2**64 - 1, size=n, dtype=np.uint64)
        gk, _ = _run(gpu, keys, None, 0, 64)
        assert np.array_equal(gk, np.sort(keys))


def test_sort_32m_u64_properties(gpu):
    """Full-size standalone sort (2^25 uniform u64 keys): sorted, and a
    permutation of the input (checksums of the multiset)."""
    n = 1 << 25
    g = torch.Generator(device="cuda").manual_seed(0)
    k = torch.randint(-(2**63), 2**63 - 1, (n,), dtype=torch.int64, device="cuda", generator=g)
    before = k.cpu().numpy().view(np.uint64).copy()
    kt = torch.empty_like(k)
    gpu.OnesweepSort().sort(k, None, kt, None, n, 0, 64)
    torch.cuda.synchronize()
    after = k.cpu().numpy().view(np.uint64)
    assert np.all(after[1:] >= after[:-1])
    assert np.array_equal(after, np.sort(before))


@pytest.mark.parametrize("cfg", [0, 1, 2, 3, 4])
def test_every_tile_config_is_bit_exact(gpu, orc, cfg):
    """The autotuner may pick any tile shape; all must give the same result."""
    print(f"sort config {cfg}", flush=True)
    gpu.set_tuning(cfg, 8, 64)
    try:
        for dtype, bits in [(np.uint32, 32), (np.uint64, 64)]:
            for n in [1, 5000, 300_007]:
                r = np.random.default_rng(cfg * 7 + n)
                keys = r.integers(0, np.iinfo(dtype).max, size=n, dtype=dtype, endpoint=True)
                keys[::9] = keys[0]
                vals = np.arange(n, dtype=np.uint32)
                for start, end in [(0, bits), (4, 28)]:
                    gk, gv = _run(gpu, keys, vals, start, end)
                    ok, ov, _ = orc.radix_sort(keys, vals, start, end)
                    assert np.array_equal(gk, ok) and np.array_equal(gv, ov), (dtype, n, start, end)
                    gk, _ = _run(gpu, keys, None, start, end)
                    assert np.array_equal(gk, ok)
    finally:
        gpu.set_tuning(2, 8, 64)


def _prefix_keys(r, n, dtype, prefix_of, shift):
    """Random keys whose 16-bit field at `shift` is prefix_of (n,) values."""
    bits = np.dtype(dtype).itemsize * 8
    keys = r.integers(0, np.iinfo(dtype).max, size=n, dtype=dtype, endpoint=True)
    mask = dtype(0xFFFF) << dtype(shift)
    keys = (keys & ~mask) | (prefix_of.astype(dtype) << dtype(shift))
    assert shift + 16 <= bits
    return keys


@pytest.mark.parametrize("case", ["tiny_segments", "one_segment", "half_one_segment", "seg_2049",
                                  "seg_3000", "seg_4097", "all_equal", "seg24_9", "seg24_12", "seg24_40",
                                  "seg24_130", "seg24_mixed", "seg24_rows"])
def test_hybrid_segment_shapes(gpu, orc, case):
    """Large sorts finish each run of equal 24-bit prefix (a segment) in LDS:
    segments of up to 9 keys by a fixed count of comparisons, longer ones
    (also ones spanning rows of 64 positions with no segment start) by a walk;
    segments longer than a block's look-ahead switch the whole sort to the
    full pass sequence.  The seg_* cases group keys by their top 16 bits
    (bits 40-47 random, so 24-bit segments of about length/256 keys); seg24_L
    makes 24-bit segments of exactly L keys, seg24_mixed of 1 to 200, and
    seg24_rows segments of 200, 200 and 112 keys that tile every 512 sorted
    positions (no segment crosses a 512-aligned boundary, so the finish needs
    no fallback, and whole rows of 64 positions hold no segment start), with
    only 8 distinct low values per segment (many ties).  Every shape must give
    the oracle's stable result bit for bit."""
    n = (1 << 20) + 17
    print(f"hybrid sort case {case} n={n}", flush=True)
    r = np.random.default_rng(sum(map(ord, case)))
    if case == "seg24_rows":
        lengths = np.resize(np.array([200, 200, 112]), 3 * (n // 512))
        group = np.concatenate([np.repeat(np.arange(lengths.size), lengths),
                                np.full(n - int(lengths.sum()), lengths.size)])
        keys = (group.astype(np.uint64) << np.uint64(40)) | r.integers(0, 8, size=n, dtype=np.uint64)
        keys = keys[r.permutation(n)]
        pf = None
    elif case.startswith("seg24_"):
        if case == "seg24_mixed":
            lengths = r.integers(1, 201, size=n // 50)
            group = np.repeat(np.arange(lengths.size), lengths)[:n]
            group = np.concatenate([group, np.arange(n - group.size) + lengths.size])
        else:
            group = np.arange(n) // int(case.split("_")[1])
        pf24 = r.permutation(int(group.max()) + 1)[group].astype(np.uint64)
        keys = (pf24 << np.uint64(40)) | (r.integers(0, 1 << 40, size=n, dtype=np.uint64))
        keys = keys[r.permutation(n)]
        keys[::7] = keys[3]  # ties keep input order
        pf = None
    elif case == "tiny_segments":
        pf = r.integers(0, 1 << 16, size=n)
    elif case == "one_segment":
        pf = np.full(n, 0x1234)
    elif case == "half_one_segment":
        pf = np.where(r.random(n) < 0.5, 0x8000, r.integers(0, 1 << 16, size=n))
    elif case == "all_equal":
        pf = None
    else:
        length = int(case.split("_")[1])
        pf = r.permutation(np.arange(n) // length)
    if case.startswith("seg24_"):
        pass
    elif pf is None:
        keys = np.full(n, 0x0123456789ABCDEF, np.uint64)
    else:
        keys = _prefix_keys(r, n, np.uint64, pf, 48)
        keys[::7] = keys[3]  # ties keep input order
    vals = np.arange(n, dtype=np.uint32)
    gk, gv = _run(gpu, keys, vals, 0, 64)
    ok, ov, _ = orc.radix_sort(keys, vals, 0, 64)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)
    gk, _ = _run(gpu, keys, None, 0, 64)
    assert np.array_equal(gk, ok)


@pytest.mark.parametrize("dtype,start,end", [(np.uint64, 4, 60), (np.uint64, 17, 41), (np.uint32, 3, 29),
                                             (np.uint32, 0, 32), (np.uint64, 40, 64), (np.uint64, 4, 64),
                                             (np.uint64, 1, 64), (np.uint64, 9, 64), (np.uint64, 7, 63),
                                             (np.uint32, 7, 32)])
def test_hybrid_bit_ranges(gpu, orc, dtype, start, end):
    """The hybrid's 24-bit prefix is the three highest digits of the requested
    range (digits past endBit included, as in every pass).  Ranges whose
    start is not a multiple of 8 and whose top digit reaches bit 63 give the
    finish kernel 40 sort bits below the prefix: the widest order word."""
    n = 600_011
    print(f"hybrid sort {np.dtype(dtype).name} [{start},{end}) n={n}", flush=True)
    r = np.random.default_rng(start * 100 + end)
    keys = r.integers(0, np.iinfo(dtype).max, size=n, dtype=dtype, endpoint=True)
    keys[::11] = keys[5]
    vals = r.integers(0, 2**32, size=n, dtype=np.uint32)
    gk, gv = _run(gpu, keys, vals, start, end)
    ok, ov, _ = orc.radix_sort(keys, vals, start, end)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gv, ov)
