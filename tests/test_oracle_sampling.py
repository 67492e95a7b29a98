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


def _unit_normals(n, seed):
    r = np.random.default_rng(seed)
    v = r.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    special = [(0, 0, 1), (0, 0, -1), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (0.6, 0, 0.8), (0, 0.6, -0.8),
               (1e-4, 0, -1.0), (0, 0, 0), (0, 0, 0)]
    v = np.vstack([v, special]).astype(np.float32)
    v[-2] = (0.0, 1.0, -0.0)  # n.z = -0: sign -1
    v[-1] = (1.0, 0.0, 0.0)
    v = (v / np.linalg.norm(v.astype(np.float64), axis=1, keepdims=True)).astype(np.float32)
    return v


def test_onb_is_orthonormal_and_right_handed(orc):
    """GetOrthonormalBasis (Duff et al. 2017, main_gpu.cu:39-45) on random
    unit normals and the poles, axes and n.z = +-0: x.n = y.n = x.y = 0,
    |x| = |y| = 1 and x cross y = n, to float rounding."""
    for n in _unit_normals(2000, 5):
        x, y = orc.onb(n)
        x, y, n64 = x.astype(np.float64), y.astype(np.float64), n.astype(np.float64)
        assert abs(x @ n64) < 4e-7 and abs(y @ n64) < 4e-7 and abs(x @ y) < 4e-7, n
        assert abs(np.linalg.norm(x) - 1) < 4e-7 and abs(np.linalg.norm(y) - 1) < 4e-7, n
        assert np.abs(np.cross(x, y) - n64).max() < 6e-7, n
    # the pole cases in closed form
    x, y = orc.onb(np.array([0, 0, 1], np.float32))
    assert x.tolist() == [1, 0, 0] and y.tolist() == [0, 1, 0]
    x, y = orc.onb(np.array([0, 0, -1], np.float32))
    assert x.tolist() == [1, 0, 0] and y.tolist() == [0, -1, 0]


def test_cosine_directions_around_a_normal(orc):
    """The AO / PT bounce direction (main_gpu.cu:152-153, :231-235): for any
    unit normal the direction is unit length, in the normal's hemisphere
    (d.n = sqrt(1 - xi1)), and over a midpoint grid of (xi0, xi1) it has the
    cosine distribution's moments E[d.n] = 2/3, E[d] = (2/3) n, and
    P(d.n >= c) = 1 - c^2."""
    g = (np.arange(48) + 0.5) / 48
    for n in _unit_normals(12, 6):
        d = np.array([orc.cosine_dir(n, a, b) for a in g for b in g], np.float64)
        n64 = n.astype(np.float64)
        assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 1e-6
        dn = d @ n64
        assert dn.min() >= -1e-7
        want = np.sqrt(1 - np.repeat(g[None, :], 48, axis=0).ravel())
        assert np.abs(dn - want).max() < 2e-6
        assert abs(dn.mean() - 2 / 3) < 2e-4
        assert np.abs(d.mean(axis=0) - (2 / 3) * n64).max() < 2e-4
        for c in (0.25, 0.5, 0.8):
            assert abs((dn >= c).mean() - (1 - c * c)) < 0.03


def _quad(y, half, up):
    """Square at height y in the xz plane; `up` picks the winding, so the
    geometric normal (v1-v0)x(v2-v1) points +y or -y."""
    a, b, c, d = (-half, y, -half), (half, y, -half), (half, y, half), (-half, y, half)
    tris = [(a, d, c), (a, c, b)] if up else [(a, b, c), (a, c, d)]
    return np.asarray(tris, np.float32)


def _render_ao(orc, tris, cam, w, h, n_sample, radius):
    nodes, root, _, _ = orc.build(tris)
    eye, ctr, up, fov = cam
    g = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    px, ao, _ = orc.render_ao(g, w, h, root, nodes, tris, 1, n_sample, radius, 0, primary_stack=True)
    return px, ao


@pytest.mark.parametrize("n_sample", [4, 16])
@pytest.mark.parametrize("winding", ["as_built", "reversed"])
def test_ao_closed_box_is_fully_occluded(orc, n_sample, winding):
    """Inside the closed cube [-1, 1]^3 every primary ray hits and, with a
    radius above the cube's diagonal, every AO ray hits a wall: ao = N * 1/N
    = 1 exactly and the pixel is black, whichever way the walls are wound.
    With the walls' normals pointing out of the box this fails if the normal
    were not turned toward the ray (the AO origin would sit 0.001 outside the
    box and every ray would leave it)."""
    from minimumlbvh_amd import scenes

    tris, _ = scenes.furnace_box()
    if winding == "reversed":
        tris = np.ascontiguousarray(tris[:, ::-1])
    px, ao = _render_ao(orc, tris, scenes.CAMERA_FURNACE, 40, 32, n_sample, 4.0)
    assert np.all(ao == 1.0) and np.all(px == 0)


@pytest.mark.parametrize("up", [True, False])
def test_ao_lone_quad_is_unoccluded(orc, up):
    """A lone square seen from above: every AO ray leaves through the
    normal's hemisphere and hits nothing, so ao = 0 and the pixel is white,
    whichever way the square is wound.  A direction built on the wrong side
    of the surface would hit the square itself."""
    tris = _quad(0.0, 1.0, up)
    cam = ((0.0, 1.0, 0.0), (0.0, 0.0, 0.001), (0.0, 0.0, 1.0), 60.0)
    px, ao = _render_ao(orc, tris, cam, 48, 48, 16, 0.5)
    hit = ao >= 0
    assert hit.sum() > 1000
    assert np.all(ao[hit] == 0.0) and np.all(px[hit] == 0xFFFFFF) and np.all(px[~hit] == 0)


def test_ao_between_parallel_planes_closed_form(orc):
    """Floor y = 0 and ceiling y = h, seen from between them: an AO ray from
    the floor (origin lifted by 0.001 along the normal, length = radius)
    reaches the ceiling iff cos(theta) >= c = (h - 0.001) / radius, which
    under the cosine distribution has probability 1 - c^2.  With h = 0.101
    and radius 0.2, E[ao] = 0.75; without the lift it would be 0.745, with
    the lift reversed 0.7399.  The pixel is the 8-bit level of
    (1 - ao)^(1/2.2) (float64 power, independent of the kernels' pow)."""
    h, radius = 0.101, 0.2
    tris = np.concatenate([_quad(0.0, 50.0, True), _quad(h, 50.0, False)])
    cam = ((0.0, h / 2, 0.0), (0.0, 0.0, 0.001), (0.0, 0.0, 1.0), 60.0)
    px, ao = _render_ao(orc, tris, cam, 128, 128, 16, radius)
    assert np.all(ao >= 0)
    c = (h - 0.001) / radius
    assert abs(ao.mean() - (1 - c * c)) < 0.0025, ao.mean()
    want = np.clip(np.trunc(np.power(1.0 - ao.astype(np.float64), 1 / 2.2) * 255 + 0.5), 0, 255).astype(np.uint32)
    assert np.array_equal(px & 0xFF, want) and np.array_equal(px >> 8 & 0xFF, want) and np.array_equal(px >> 16, want)


@pytest.mark.parametrize("up", [True, False])
def test_normal_image_of_an_axis_aligned_quad(orc, up):
    """main_gpu.cu:84-109: the pixel is packRGBA((n + 1) / 2) of the
    unflipped unit geometric normal: (0, +-1, 0) gives (128, 255, 128) or
    (128, 0, 128); misses are black."""
    tris = _quad(0.0, 1.0, up)
    nodes, root, _, _ = orc.build(tris)
    g = orc.raygen_lookat((0.0, 1.0, 0.0), (0.0, 0.0, 0.001), (0.0, 0.0, 1.0), math.radians(60), 40, 40)
    ng = np.cross(tris[0, 1] - tris[0, 0], tris[0, 2] - tris[0, 1]).astype(np.float64)
    n = ng / np.linalg.norm(ng)
    assert np.allclose(n, [0, 1 if up else -1, 0])
    rgb = np.trunc((n + 1) / 2 * 255 + 0.5).astype(np.uint32)
    want = rgb[2] << 16 | rgb[1] << 8 | rgb[0]
    px = orc.render_normal(g, 40, 40, root, nodes, tris)
    assert (px == want).sum() > 1000 and np.all((px == want) | (px == 0))


def _planes_scene(h, half, floor_albedo, light):
    """Diffuse floor at y = 0 and an emissive, black ceiling at y = h."""
    tris = np.concatenate([_quad(0.0, half, True), _quad(h, half, True)])
    attribs = np.zeros((4, 15), np.float32)
    attribs[:2, 9:12] = floor_albedo
    attribs[2:, 12:15] = light
    return tris, attribs


@pytest.mark.parametrize("up", [True, False])
def test_pt_one_bounce_to_an_emissive_ceiling(orc, up):
    """Path tracing (main_gpu.cu:179-271) with max depth 2 under an infinite
    emissive ceiling: the floor emits nothing, its bounce (p = 1 for albedo
    0.5) always reaches the ceiling, so every path carries exactly
    albedo * Le.  A bounce into the floor's lower hemisphere, or an origin
    offset to the wrong side, would miss and give 0."""
    light = np.array([1.0, 2.0, 3.0], np.float32)
    tris, attribs = _planes_scene(0.05, 1000.0, 0.5, light)
    if not up:
        tris[:2] = tris[:2, ::-1]
    nodes, root, _, _ = orc.build(tris)
    g = orc.raygen_lookat((0.0, 0.025, 0.0), (0.0, 0.0, 0.001), (0.0, 0.0, 1.0), math.radians(60), 24, 24)
    acc = np.zeros((24 * 24, 4), np.float32)
    for _ in range(4):
        acc = orc.render_pt(acc, g, 24, 24, root, nodes, tris, attribs, 1, 1024, 2)
    assert np.array_equal(acc, np.tile(np.append(4 * 0.5 * light, 4).astype(np.float32), (24 * 24, 1)))


def _square_form_factor(a, h):
    """Cosine-weighted fraction of directions from a point that hit a
    parallel square of side a centred at height h above it: four corner
    rectangles of the differential-area-to-parallel-rectangle form factor."""
    A = B = (a / 2) / h

    def corner(A, B):
        sa, sb = math.sqrt(1 + A * A), math.sqrt(1 + B * B)
        return (A / sa * math.atan(B / sa) + B / sb * math.atan(A / sb)) / (2 * math.pi)

    return 4 * corner(A, B)


def test_pt_direct_light_form_factor(orc):
    """A floor point under a square light of side a at height h receives
    albedo * Le * F with F the point-to-square form factor: the path
    tracer's bounce directions have the cosine distribution.  The camera
    sees a 0.004-wide patch of floor under the light's centre."""
    h, a = 0.2, 0.3
    F = _square_form_factor(a, h)
    floor = _quad(0.0, 50.0, True)
    light = _quad(h, a / 2, False)
    tris = np.concatenate([floor, light])
    attribs = np.zeros((4, 15), np.float32)
    attribs[:2, 9:12] = 0.5
    attribs[2:, 12:15] = 1.0
    nodes, root, _, _ = orc.build(tris)
    w = hh = 64
    g = orc.raygen_lookat((0.0, 0.1, 0.0), (0.0, 0.0, 1e-4), (0.0, 0.0, 1.0), math.radians(2.3), w, hh)
    acc = np.zeros((w * hh, 4), np.float32)
    spp = 16
    for _ in range(spp):
        acc = orc.render_pt(acc, g, w, hh, root, nodes, tris, attribs, 1, 1024, 2)
    mean = acc[:, 0].sum() / acc[:, 3].sum()
    assert abs(mean - 0.5 * F) < 0.004, (mean, 0.5 * F)
