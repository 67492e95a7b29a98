"""The C++ facade as a drop-in: a reference-style application
(tests/facade/facade_demo.hip: OnesweepSort + BVHGPUBuilder + a user kernel
calling intersect_stackfree, the sort self-test of main_gpu.cpp:104-145, the
reference's per-leaf build kernel calling build_lbvh on the device, and a
user kernel running intersect_stackfree / intersect_stackfull /
intersect_escape_link) built against include/minimum_lbvh.h only.  Its node
arrays, roots, normal image and Hit records must be bit-identical to the CPU
oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from minimumlbvh_amd import _build, scenes
from tests.test_gpu_trace import random_rays

pytestmark = pytest.mark.gpu


def test_facade_application_matches_oracle(orc, tmp_path):
    assert os.path.exists(_build.FACADE_BIN), "build the facade demo first (__graft_entry__.build())"
    tris = scenes.random_triangles(65536, 0)
    tris.tofile(tmp_path / "tris.bin")
    w, h = 200, 160
    eye, ctr, up, fov = scenes.CAMERA_SMALL
    args = [str(x) for x in (*eye, *ctr, *up, math.radians(fov))]
    rays = random_rays(20_000, 11)
    rays[5000:10000, 3] = 0.05
    rays.tofile(tmp_path / "rays.bin")
    on, oroot, _, _ = orc.build(tris)
    depth = _depth(on, oroot)
    print(f"facade demo: oracle tree depth {depth}", flush=True)
    assert depth < 64, "the demo's stackfull kernel holds 64 entries"
    r = subprocess.run([_build.FACADE_BIN, str(tmp_path / "tris.bin"), str(tmp_path / "out"), str(w), str(h), *args,
                        str(tmp_path / "rays.bin")], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "sort self-test N=1048579: OK" in r.stdout
    nodes = np.fromfile(tmp_path / "out_nodes.bin", np.uint32).reshape(-1, 16)
    root = int(np.fromfile(tmp_path / "out_root.bin", np.uint32)[0])
    px = np.fromfile(tmp_path / "out_pixels.bin", np.uint32)
    assert root == oroot and np.array_equal(nodes, on)
    # the reference's per-leaf build kernel with the facade's build_lbvh
    lnodes = np.fromfile(tmp_path / "out_leafbuild_nodes.bin", np.uint32).reshape(-1, 16)
    lroot = int(np.fromfile(tmp_path / "out_leafbuild_root.bin", np.uint32)[0])
    assert lroot == oroot and np.array_equal(lnodes, on)
    # intersect_stackfree closest / any, intersect_stackfull, intersect_escape_link in a user kernel
    hits = np.fromfile(tmp_path / "out_hits.bin", np.uint32).reshape(-1, 4, 8)
    for col, (query, method) in enumerate([(0, 0), (1, 0), (0, 1), (0, 2)]):
        oh, _ = orc.trace(on, tris, oroot, rays, query=query, method=method)
        oh = oh.view(np.uint32).reshape(-1, 8)
        assert np.array_equal(hits[:, col, [0, 2, 3, 4, 5, 6, 7]], oh[:, [0, 2, 3, 4, 5, 6, 7]]), col
    assert (hits[:, 0, 0].view(np.float32) < 3e38).sum() > 2000
    og = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    ref = orc.render_normal(og, w, h, oroot, on, tris)
    assert (ref != 0).sum() > 500
    assert np.array_equal(px, ref)


def _depth(nodes, root):
    """Internal nodes on the deepest root-to-leaf path."""
    depth, st = 0, [(root, 1)]
    while st:
        i, d = st.pop()
        depth = max(depth, d)
        for c in nodes[i, 0:2]:
            if not c & 0x80000000:
                st.append((int(c), d + 1))
    return depth
