"""Pins the oracle's build pipeline (minimum_lbvh.h:284-361, 512-592):
known answer, numpy restatement of keys/sort/deltas, and the Cartesian-tree
characterisation of the topology and AABBs (tests/lbvh_checks.py)."""
import numpy as np
import pytest

from minimumlbvh_amd import scenes
from tests import lbvh_checks as C


def test_toy_tree_known_answer(orc):
    """main.cpp:60-109: keys {2,3,4,5,8,12,13,15}, no triangles."""
    keys = np.array([0b0010, 0b0011, 0b0100, 0b0101, 0b1000, 0b1100, 0b1101, 0b1111], np.uint32)
    d = np.array([orc.lib().orc_delta32(int(keys[i]), int(keys[i + 1])) for i in range(7)], np.uint8)
    assert d.tolist() == [1, 3, 1, 4, 3, 1, 2]
    nodes, root = orc.build_from_deltas(d)
    L = 0x80000000
    INV = 0x7FFFFFFF
    F = 0xFFFFFFFF
    expect = {  # node: (child0, child1, parent, context)
        0: (L | 0, L | 1, 1, 1),
        1: (0, 2, 3, 3),
        2: (L | 2, L | 3, 1, 3),
        3: (1, 4, INV, F),
        4: (L | 4, 6, 3, F),
        5: (L | 5, L | 6, 6, 6),
        6: (5, L | 7, 4, F),
    }
    assert root == 3
    for i, (c0, c1, p, ctx) in expect.items():
        assert tuple(int(x) for x in nodes[i, [0, 1, 14, 15]]) == (c0, c1, p, ctx), i
    assert orc.validate(root, nodes, d, 2**30)
    # the independent characterisation agrees on the toy keys as well
    tris = np.zeros((8, 3, 3), np.float32)
    exp, eroot = C.expected_tree(tris, np.arange(8, dtype=np.uint32), d)
    assert eroot == root
    assert np.array_equal(exp[:, [0, 1, 14, 15]], nodes[:, [0, 1, 14, 15]])


def test_scene_aabb_matches_numpy(orc):
    tris = scenes.random_triangles(5000, 7)
    box = orc.scene_aabb(tris)
    v = tris.reshape(-1, 3)
    assert np.array_equal(box[:3], v.min(0))
    assert np.array_equal(box[3:], v.max(0))


@pytest.mark.parametrize("use_normal", [False, True])
def test_morton_keys_match_numpy(orc, use_normal):
    tris = scenes.random_triangles(20000, 5)
    rec = orc.morton_records(tris, opt=1 if use_normal else 0)
    assert np.array_equal((rec >> np.uint64(32)).astype(np.uint32), np.arange(20000, dtype=np.uint32))
    keys = (rec & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(keys, C.numpy_morton_keys(tris, use_normal))


def test_morton_key_decodes_to_quantised_centroid(orc):
    """Decoding the 63-bit code recovers the quantised centroid; the stored key
    is its top 32 bits."""
    tris = scenes.random_triangles(3000, 9)
    keys = (orc.morton_records(tris) & np.uint64(0xFFFFFFFF)).astype(np.uint64)
    lo = tris.reshape(-1, 3).min(0)
    hi = tris.reshape(-1, 3).max(0)
    c = ((tris[:, 0] + tris[:, 1]) + tris[:, 2]) / np.float32(3)
    q = np.clip(np.trunc(((c - lo) / (hi - lo)) * np.float32(2097152)), 0, 2097151).astype(np.uint32)
    xyz = orc.decode_batch(keys << np.uint64(31))
    # the low 31 bits of the code are dropped: compare the surviving high bits
    # (x keeps bits 11..20, y and z keep 10..20 / 11..20 depending on phase)
    code = orc.morton_batch(q)
    assert np.array_equal(code >> np.uint64(31), keys)
    assert np.array_equal(xyz >> 11, q >> 11)


def test_sort_is_stable_argsort(orc):
    tris = scenes.random_triangles(30000, 2)
    tris[1000:1100] = tris[5]  # force equal keys
    rec = orc.morton_records(tris)
    keys = (rec & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    srt = rec.copy()
    orc.lib().orc_sort_records(srt.ctypes.data_as(orc.c_void_p), srt.shape[0])
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(srt, rec[order])


def test_deltas_closed_form(orc):
    tris = scenes.random_triangles(10000, 4)
    nodes, root, rec, d = orc.build(tris)
    sk = (rec & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    assert np.array_equal(d, C.deltas_of(sk))


def _check_against_characterisation(orc, tris, opt=0):
    n = tris.shape[0]
    nodes, root, rec, d = orc.build(tris, opt=opt)
    keys = C.numpy_morton_keys(tris, bool(opt & 1))
    order = C.stable_order(keys)
    assert np.array_equal((rec >> np.uint64(32)).astype(np.uint32), order)
    assert np.array_equal(d, C.deltas_of(keys[order]))
    exp, eroot = C.expected_tree(tris, order, d)
    assert root == eroot
    assert np.array_equal(nodes, exp)
    C.structural_checks(nodes, root, n)
    assert orc.validate(root, nodes, d)
    return nodes, root, d


@pytest.mark.parametrize("n", [2, 3, 4, 257, 4096, 65536])
def test_build_matches_characterisation(orc, n):
    tris = scenes.random_triangles(n, 100 + n)
    _check_against_characterisation(orc, tris)


def test_build_use_normal(orc):
    _check_against_characterisation(orc, scenes.random_triangles(4096, 77), opt=1)


def test_build_degenerate_inputs(orc):
    # all centroids identical -> all keys equal -> an (n-1)-deep chain
    t = np.tile(scenes.random_triangles(1, 1), (300, 1, 1))
    nodes, root, d = _check_against_characterisation(orc, t)
    assert d.max() == 0 and root == 0
    # flat scene axis (all z equal): extent 0 maps to coordinate 0
    t = scenes.random_triangles(1000, 3)
    t[:, :, 2] = 0.25
    _check_against_characterisation(orc, t)
    # many exact duplicates mixed with distinct triangles
    t = scenes.random_triangles(2000, 8)
    t[::3] = t[1]
    _check_against_characterisation(orc, t)


def test_build_signed_zero_coordinates(orc):
    """Vertices on the coordinate planes written as -0 or +0 (the reference
    Cornell box's OBJ has both): node boxes take -0 as the lower and +0 as
    the upper bound wherever a merge meets both, so the serial climb, the
    parallel one and the characterisation give the same bits."""
    r = np.random.default_rng(41)
    t = scenes.random_triangles(20000, 41)
    snap = r.random(t.shape) < 0.3
    t[snap] = np.where(r.random(int(snap.sum())) < 0.5, np.float32(-0.0), np.float32(0.0))
    nodes, root, d = _check_against_characterisation(orc, t)
    assert (nodes[:, 2:14] == 0x80000000).any() and (nodes[:, 2:14] == 0).any()
    for threads in (1, 3, 8):
        b = orc.build(t, parallel=True, nthreads=threads)
        assert np.array_equal(b[0], nodes) and b[1] == root


def test_comparison_sort_build_reference_structure(orc):
    """The CPU-baseline build (reference structure: std::sort on the key
    alone, minimum_lbvh.h:554) gives a valid LBVH over the same sorted keys;
    with all keys distinct the order cannot differ from the stable sort, so
    the tree is the radix build's, bit for bit.  With ties, only the order of
    equal-key leaves may differ."""
    t = scenes.random_triangles(65536, 0)
    a = orc.build(t, opt=0)
    b = orc.build(t, opt=8)
    keys = a[2] & 0xFFFFFFFF
    assert np.unique(keys).size == keys.size
    assert np.array_equal(a[0], b[0]) and a[1] == b[1] and np.array_equal(a[2], b[2])
    t = scenes.random_triangles(20000, 4)
    t[::3] = t[5]  # many equal keys
    a = orc.build(t, opt=0)
    b = orc.build(t, opt=8, parallel=False)
    assert np.array_equal(a[2] & 0xFFFFFFFF, b[2] & 0xFFFFFFFF)
    assert np.array_equal(np.sort(a[2]), np.sort(b[2]))
    assert np.array_equal(a[3], b[3])
    C.structural_checks(b[0], b[1], t.shape[0])
    assert orc.validate(b[1], b[0], b[3])


def test_build_serial_equals_parallel(orc):
    tris = scenes.random_triangles(65536, 0)
    a = orc.build(tris, parallel=True)
    b = orc.build(tris, parallel=False)
    assert np.array_equal(a[0], b[0]) and a[1] == b[1]


def test_build_rejects_small_n(orc):
    with pytest.raises(ValueError):
        orc.build(scenes.random_triangles(1, 0))


def test_validate_detects_violation(orc):
    tris = scenes.random_triangles(512, 1)
    nodes, root, rec, d = orc.build(tris)
    bad = d.copy()
    bad[root] = 0  # the root must hold the largest delta
    assert not orc.validate(root, nodes, bad)


MORTON64 = 4


def test_morton64_build_matches_characterisation(orc):
    """63-bit keys on a clustered scene (many triangles share a 32-bit key):
    sorted order = stable argsort of the numpy 63-bit codes, deltas = their
    bit lengths, topology and boxes = the Cartesian-tree characterisation."""
    tris = scenes.random_triangles(20000, 30, sigma=0.0005)
    tris[:12000] = tris[:12000] * np.float32(0.002) + np.float32(0.5)  # a dense clump
    keys32 = C.numpy_morton_keys(tris)
    assert np.unique(keys32).size < 0.7 * tris.shape[0]
    nodes, root, rec, d = orc.build(tris, opt=MORTON64)
    keys = C.numpy_morton_keys(tris, full=True)
    order = C.stable_order(keys)
    assert np.array_equal((rec >> np.uint64(32)).astype(np.uint32), order)
    assert np.array_equal((rec & np.uint64(0xFFFFFFFF)), keys[order] >> np.uint64(31))
    assert np.array_equal(d, C.deltas_of(keys[order]))
    assert d.max() <= 63 and np.unique(keys).size > np.unique(keys32).size
    exp, eroot = C.expected_tree(tris, order, d)
    assert root == eroot and np.array_equal(nodes, exp)
    C.structural_checks(nodes, root, tris.shape[0])
    assert orc.validate(root, nodes, d, 64)


def test_morton64_equals_32bit_build_when_keys_are_distinct(orc):
    """With every 32-bit key distinct the two variants sort alike and each
    63-bit delta is the 32-bit delta + 31 (the highest differing bit lies in
    the top 32 bits), a monotone shift: the climb takes the same decisions and
    the node arrays are identical."""
    tris = scenes.random_triangles(3000, 5)
    keys32 = C.numpy_morton_keys(tris)
    assert np.unique(keys32).size == keys32.size
    a_nodes, a_root, a_rec, a_d = orc.build(tris)
    b_nodes, b_root, b_rec, b_d = orc.build(tris, opt=MORTON64)
    assert np.array_equal(b_rec, a_rec)
    assert np.array_equal(b_d.astype(int), a_d.astype(int) + 31)
    assert b_root == a_root and np.array_equal(b_nodes, a_nodes)


def test_morton64_rejects_use_normal(orc):
    with pytest.raises(ValueError):
        orc.build(scenes.random_triangles(100, 1), opt=MORTON64 | 1)
