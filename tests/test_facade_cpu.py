"""The C++ facade's host side (include/minimum_lbvh.h compiled into
tests/facade/facade_cpu.cpp), no GPU: BVHCPUBuilder::build (minimum_lbvh.h:
509-669) and validate(), AABB::encodeMortonCode, hashPCG, and the three
traversals intersect_stackfree / intersect_stackfull / intersect_escape_link
(minimum_lbvh.h:912-1137) called on the host, each compared bit for bit with
the CPU oracle.  validate() must abort on a tree that breaks its invariant."""
import os
import subprocess

import numpy as np
import pytest

from minimumlbvh_amd import _build, scenes
from tests.test_gpu_trace import random_rays

HIT = np.dtype([("t", "<f4"), ("pad", "<u4"), ("u", "<f4"), ("v", "<f4"), ("ng", "<f4", (3,)), ("tri", "<u4")])


def _bin():
    if not os.path.exists(_build.FACADE_CPU_BIN):
        pytest.skip("facade CPU program not built (run __graft_entry__.build())")
    return _build.FACADE_CPU_BIN


def _run_build(tmp_path, tris, rays, opt):
    (tmp_path / "t.bin").write_bytes(np.ascontiguousarray(tris, np.float32).tobytes())
    (tmp_path / "r.bin").write_bytes(np.ascontiguousarray(rays, np.float32).tobytes())
    r = subprocess.run([_bin(), "build", str(tmp_path / "t.bin"), str(tmp_path / "r.bin"), str(tmp_path / "o"),
                        str(opt)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    nodes = np.fromfile(tmp_path / "o_nodes.bin", np.uint32).reshape(-1, 16)
    root = int(np.fromfile(tmp_path / "o_root.bin", np.uint32)[0])
    deltas = np.fromfile(tmp_path / "o_deltas.bin", np.uint8)
    hits = np.fromfile(tmp_path / "o_hits.bin", HIT).reshape(-1, 4)
    return nodes, root, deltas, hits


def _hits_equal(a, b):
    """Every field of the 32-byte Hit records but the padding word."""
    a = np.ascontiguousarray(a).view(np.uint32).reshape(-1, 8)
    b = np.ascontiguousarray(b).view(np.uint32).reshape(-1, 8)
    return np.array_equal(a[:, [0, 2, 3, 4, 5, 6, 7]], b[:, [0, 2, 3, 4, 5, 6, 7]])


@pytest.mark.parametrize("case,opt", [("random", 0), ("random", 1), ("random", 2), ("random", 3), ("ties", 2),
                                      ("signed_zero", 2)])
def test_cpu_builder_and_host_traversals_match_oracle(orc, tmp_path, case, opt):
    tris = scenes.random_triangles(30_000 if case != "random" else 65_536, 3)
    if case == "ties":
        tris[::4] = tris[9]
    elif case == "signed_zero":
        r = np.random.default_rng(2)
        snap = r.random(tris.shape) < 0.3
        tris[snap] = np.where(r.random(int(snap.sum())) < 0.5, np.float32(-0.0), np.float32(0.0))
    rays = random_rays(3000, 5)
    rays[1000:2000, 3] = 0.05  # short rays: any-hit and closest differ more often
    nodes, root, deltas, hits = _run_build(tmp_path, tris, rays, opt)
    on, oroot, _, od = orc.build(tris, opt=opt & 1, parallel=bool(opt & 2))
    assert root == oroot and np.array_equal(nodes, on) and np.array_equal(deltas, od)
    for col, (query, method) in enumerate([(0, 0), (1, 0), (0, 1), (0, 2)]):
        oh, _ = orc.trace(on, tris, oroot, rays, query=query, method=method)
        assert _hits_equal(hits[:, col], oh), (col, query, method)
    assert (hits[:, 0]["t"] < 3e38).sum() > 500


def test_validate_aborts_on_a_broken_invariant(tmp_path):
    tris = scenes.random_triangles(2000, 8)
    (tmp_path / "t.bin").write_bytes(tris.tobytes())
    r = subprocess.run([_bin(), "corrupt", str(tmp_path / "t.bin")], capture_output=True, text=True, timeout=60)
    assert r.returncode in (-6, 134), (r.returncode, r.stdout)
    assert "did not abort" not in r.stdout


def test_encode_morton_and_hash_pcg_match_oracle(orc, tmp_path):
    pts = np.random.default_rng(4).normal(size=(50_000, 3)).astype(np.float32)
    pts[:10] = [[0, 0, 0], [-0.0, 0, 0], [1e-30, 0, 0], [5, 5, 5], [-5, -5, -5], [0, 5, -5], [3, 3, 3],
                [-3, 0, 2], [2, -1, 0], [0, 0, 1]]
    (tmp_path / "p.bin").write_bytes(pts.tobytes())
    r = subprocess.run([_bin(), "misc", str(tmp_path / "p.bin"), str(tmp_path / "m")], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    box = np.fromfile(tmp_path / "m_box.bin", np.float32)
    codes = np.fromfile(tmp_path / "m_codes.bin", np.uint64)
    hashes = np.fromfile(tmp_path / "m_hashes.bin", np.uint32)
    obox, ocodes = orc.points_morton(pts)
    assert np.array_equal(box.view(np.uint32), obox.view(np.uint32))
    assert np.array_equal(codes, ocodes)
    assert codes.max() < (1 << 63) and np.unique(codes).size > 40_000
    L = orc.lib()
    want = [L.orc_hash_pcg((i * 2654435761) & 0xFFFFFFFF) for i in range(4096)]
    assert hashes.tolist() == want
