"""Pins the oracle's sampler, camera and shading helpers: golden vectors
from the reference's own sobol.h / camera.h (tests/golden/, regenerated by
tests/golden/make_golden.py through oracle/_ref), the (0,m,2)-net property
of the scrambled Sobol points, and closed forms."""
import math
import os

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def test_sobol_golden(orc):
    d = np.load(os.path.join(GOLD, "sobol.npz"))
    q, xy = d["inputs"], d["outputs"]
    for (i, p0, p1, p2), (x, y) in zip(q.tolist(), xy):
        gx, gy = orc.sobol_2d(i, p0, p1, p2)
        assert gx.view(np.uint32) == x.view(np.uint32) and gy.view(np.uint32) == y.view(np.uint32)


def test_camera_golden(orc):
    d = np.load(os.path.join(GOLD, "camera.npz"))
    for rec, out in zip(d["inputs"], d["outputs"]):
        g = orc.raygen_lookat(rec[0:3], rec[3:6], rec[6:9], float(rec[9]), int(rec[10]), int(rec[11]))
        got = np.array(list(g.origin) + list(g.right) + list(g.up), np.float32)
        ro, rd = orc.raygen_shoot(g, float(rec[12]), float(rec[13]))
        got = np.concatenate([got, ro, rd])
        assert np.array_equal(got.view(np.uint32), out.view(np.uint32))


def test_live_reference_driver_matches(orc, tmp_path):
    """When oracle/_ref was built (reference sources present), compare live."""
    if not os.path.exists(orc.REF_DRIVER):
        pytest.skip("oracle/_ref not built (no reference sources here)")
    import subprocess

    rng = np.random.default_rng(99)
    q = rng.integers(0, 2**32, size=(500, 4), dtype=np.uint64).astype(np.uint32)
    inp = tmp_path / "in.bin"
    out = tmp_path / "out.bin"
    q.tofile(inp)
    subprocess.run([orc.REF_DRIVER, "sobol", str(inp), str(out)], check=True)
    ref = np.fromfile(out, np.float32).reshape(-1, 2)
    for (i, p0, p1, p2), (x, y) in zip(q.tolist(), ref):
        gx, gy = orc.sobol_2d(i, p0, p1, p2)
        assert gx == x and gy == y


@pytest.mark.parametrize("m", [4, 6, 8])
def test_sobol_is_a_0m2_net(orc, m):
    """Owen scrambling keeps the (0,m,2)-net property: the first 2^m points
    put exactly one point in every elementary box of area 2^-m."""
    n = 1 << m
    for seed in (0, 12345):
        pts = np.array([orc.sobol_2d(i, 7, 3, seed) for i in range(n)], np.float64)
        assert np.all((pts >= 0) & (pts < 1))
        for k in range(m + 1):
            a = np.floor(pts[:, 0] * (1 << k)).astype(int)
            b = np.floor(pts[:, 1] * (1 << (m - k))).astype(int)
            cells = a * (1 << (m - k)) + b
            assert np.array_equal(np.sort(cells), np.arange(n)), (seed, k)


def test_camera_centre_ray_is_forward(orc):
    g = orc.raygen_lookat((0.5, 0.5, -1.5), (0.5, 0.5, -0.5), (0, 1, 0), math.radians(60), 256, 256)
    ro, rd = orc.raygen_shoot(g, 0.5, 0.5)
    assert np.allclose(rd, [0, 0, 1], atol=1e-6)
    assert np.allclose(ro, [0.5, 0.5, -1.5])
    # right = normalize(cross(forward, up)) = -x here; u=0,v=0 is (-right, +up)
    assert np.allclose(np.array(g.right), [-math.tan(math.radians(30)), 0, 0], rtol=1e-6)
    _, corner = orc.raygen_shoot(g, 0.0, 0.0)
    assert math.isclose(corner[0], math.tan(math.radians(30)), rel_tol=1e-6)
    assert math.isclose(corner[1], math.tan(math.radians(30)), rel_tol=1e-6)


def test_pack_rgba_closed_form(orc):
    L = orc.lib()
    assert L.orc_pack_rgba(0, 0, 0, 1) == 0
    assert L.orc_pack_rgba(1, 1, 1, 1) == 0xFFFFFF
    assert L.orc_pack_rgba(2.0, -1.0, 0.5, 1) == (128 << 16) | 255  # clamp, round half up
    v = np.float32(100.4 / 255)
    assert L.orc_pack_rgba(float(v), 0, 0, 0) == 100


def test_hash_pcg_known_values(orc):
    # PCG hash of Jarzynski & Olano 2020, evaluated in Python integers
    def h(v):
        state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
        word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
        return ((word >> 22) ^ word) & 0xFFFFFFFF

    for v in [0, 1, 2, 12345, 0xFFFFFFFF, 0x80000000]:
        assert orc.lib().orc_hash_pcg(v) == h(v)
