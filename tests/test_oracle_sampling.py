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


def test_pcg32_published_known_answer(orc):
    """The PCG of main_gpu.cu:51-82 is PCG32 (XSH RR, 64-bit state, seeded
    as pcg32_srandom_r): the published pcg32-demo outputs for seed 42,
    sequence 54 pin it."""
    out = np.zeros(6, np.uint32)
    orc.lib().orc_pcg32(42, 54, 6, out.ctypes.data)
    assert out.tolist() == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_uniformf_is_the_top_23_bits(orc):
    n = 64
    u = np.zeros(n, np.uint32)
    f = np.zeros(n, np.float32)
    orc.lib().orc_pcg32(7, 0, n, u.ctypes.data)
    orc.lib().orc_pcg32_uniformf(7, 0, n, f.ctypes.data)
    assert np.array_equal(f, ((u >> 9).astype(np.float64) / 2.0**23).astype(np.float32))


def test_hemisphere_cos_points_and_moments(orc):
    """Cosine-weighted hemisphere of main_gpu.cu:24-36: xi0 sets the azimuth
    (x = r cos, z = r sin), xi1 = r^2, y = sqrt(1 - r^2).  Fixed points, unit
    length, and over a 256 x 256 midpoint grid the moments of the cosine
    distribution: E[y] = 2/3, E[x] = E[z] = 0, E[x^2] = E[z^2] = 1/4."""
    L = orc.lib()

    def hc(a, b):
        v = np.zeros(3, np.float32)
        L.orc_hemisphere_cos(a, b, v.ctypes.data)
        return v.astype(np.float64)

    assert hc(0.0, 0.0).tolist() == [0.0, 1.0, 0.0]
    assert np.allclose(hc(0.25, 1.0), [0.0, 0.0, 1.0], atol=1e-6)
    assert np.allclose(hc(0.5, 0.25), [-0.5, math.sqrt(0.75), 0.0], atol=1e-6)
    assert np.allclose(hc(0.125, 0.5), [0.5, math.sqrt(0.5), 0.5], atol=1e-6)
    assert np.allclose(hc(0.75, 0.36), [0.0, 0.8, -0.6], atol=1e-6)
    g = (np.arange(256) + 0.5) / 256
    v = np.array([hc(a, b) for a in g for b in g])
    assert np.all(np.abs(np.linalg.norm(v, axis=1) - 1) < 1e-6)
    assert np.all(v[:, 1] >= 0)
    m = v.mean(axis=0)
    assert abs(m[1] - 2 / 3) < 1e-4 and abs(m[0]) < 1e-6 and abs(m[2]) < 1e-6
    assert np.allclose((v[:, [0, 2]] ** 2).mean(axis=0), 0.25, atol=1e-5)


@pytest.mark.parametrize("sobol", [1, 0])
def test_pt_furnace_closed_form(orc, sobol):
    """Path tracing inside a closed cube whose walls all emit 1 and reflect
    0.5 (main_gpu.cu:179-252): emission is collected at hits 0 .. depth-1
    with throughput 0.5^k, Russian roulette keeps p = min(max(tp) / 0.25, 1),
    so every path's radiance is exactly 1, 1.5, 1.75 at max depth 1, 2, 3;
    at depth 4 the last hit survives with p = 0.5 and weight 0.25 / 0.5, so
    a path gives 1.75 or 2.0, about half each; depth 0 gives 0.  Every
    sample is counted in the alpha channel."""
    from minimumlbvh_amd import scenes

    tris, attribs = scenes.furnace_box()
    nodes, root, _, _ = orc.build(tris)
    eye, ctr, up, fov = scenes.CAMERA_FURNACE
    w, h, spp = 48, 40, 3
    g = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    for depth, exact in ((0, 0.0), (1, 1.0), (2, 1.5), (3, 1.75), (4, None)):
        acc = np.zeros((w * h, 4), np.float32)
        paths = []
        for _ in range(spp):
            before = acc[:, 0].copy()
            acc = orc.render_pt(acc, g, w, h, root, nodes, tris, attribs, sobol, 1024, depth)
            paths.append(acc[:, 0] - before)
        paths = np.array(paths)
        assert np.array_equal(acc[:, 3], np.full(w * h, spp, np.float32))
        assert np.array_equal(acc[:, 0], acc[:, 1]) and np.array_equal(acc[:, 0], acc[:, 2])
        if exact is not None:
            assert np.all(paths == np.float32(exact)), depth
        else:
            assert np.all((paths == np.float32(1.75)) | (paths == np.float32(2.0)))
            assert 0.45 < (paths == np.float32(2.0)).mean() < 0.55
