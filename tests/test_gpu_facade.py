"""The C++ facade as a drop-in: a reference-style application
(tests/facade/facade_demo.hip: OnesweepSort + BVHGPUBuilder + a user kernel
calling intersect_stackfree, and the sort self-test of main_gpu.cpp:104-145)
built against include/minimum_lbvh.h only.  Its node array, root and normal
image must be bit-identical to the CPU oracle."""
import math
import os
import subprocess

import numpy as np
import pytest

from minimumlbvh_amd import _build, scenes

pytestmark = pytest.mark.gpu


def test_facade_application_matches_oracle(orc, tmp_path):
    assert os.path.exists(_build.FACADE_BIN), "build the facade demo first (__graft_entry__.build())"
    tris = scenes.random_triangles(65536, 0)
    tris.tofile(tmp_path / "tris.bin")
    w, h = 200, 160
    eye, ctr, up, fov = scenes.CAMERA_SMALL
    args = [str(x) for x in (*eye, *ctr, *up, math.radians(fov))]
    r = subprocess.run([_build.FACADE_BIN, str(tmp_path / "tris.bin"), str(tmp_path / "out"), str(w), str(h), *args],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "sort self-test N=1048579: OK" in r.stdout
    nodes = np.fromfile(tmp_path / "out_nodes.bin", np.uint32).reshape(-1, 16)
    root = int(np.fromfile(tmp_path / "out_root.bin", np.uint32)[0])
    px = np.fromfile(tmp_path / "out_pixels.bin", np.uint32)
    on, oroot, _, _ = orc.build(tris)
    assert root == oroot and np.array_equal(nodes, on)
    og = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    ref = orc.render_normal(og, w, h, oroot, on, tris)
    assert (ref != 0).sum() > 500
    assert np.array_equal(px, ref)
