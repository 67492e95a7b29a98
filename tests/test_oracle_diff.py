"""Pins the fp64 adjoint of the primary-hit shading against fp64 central
finite differences of the fp64 forward (hit triangles held fixed)."""
import math

import numpy as np

from minimumlbvh_amd import scenes


def _setup(orc, level=2, w=48, h=40):
    v, f = scenes.displaced_icosphere(level, seed=1)
    tris = scenes.deindex(v, f)
    nodes, root, _, _ = orc.build(tris)
    eye, ctr, up, fov = scenes.CAMERA_SPHERE
    g = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    rays = np.zeros((w * h, 8), np.float32)
    for p in range(w * h):
        ro, rd = orc.raygen_shoot(g, (p % w) / w, (p // w) / h)
        rays[p, 0:3], rays[p, 4:7] = ro, rd
    rays[:, 3] = np.float32(3.402823466e38)
    hits, _ = orc.trace(nodes, tris, root, rays)
    return v.astype(np.float64), f, hits["tri"].copy(), g, w, h


def test_adjoint_matches_central_fd(orc):
    v, f, hit, g, w, h = _setup(orc)
    light = np.array([0.3, 0.8, -0.52])
    light /= np.linalg.norm(light)
    assert (hit != 0xFFFFFFFF).sum() > 300
    rng = np.random.default_rng(0)
    target = orc.primary_eval64(v + rng.normal(0, 0.01, v.shape), f, hit, g, w, h, light)
    wt = 0.5

    def loss(vv):
        o = orc.primary_eval64(vv, f, hit, g, w, h, light)
        return wt * np.sum((o[:, 0] - target[:, 0]) ** 2) + np.sum((o[:, 1] - target[:, 1]) ** 2)

    o = orc.primary_eval64(v, f, hit, g, w, h, light)
    gout = np.stack([2 * wt * (o[:, 0] - target[:, 0]), 2 * (o[:, 1] - target[:, 1])], axis=1)
    grad = orc.primary_grad64(v, f, hit, g, w, h, light, gout)
    touched = np.unique(f[hit[hit != 0xFFFFFFFF]].ravel())
    pick = rng.choice(touched, size=min(60, touched.size), replace=False)
    # fp64 central differences: the truncation error falls as eps^2 (grazing
    # pixels make t strongly curved: 4e-3 relative at eps 1e-4, 4e-7 at 1e-6)
    eps = 1e-6
    for j in pick:
        for c in range(3):
            vp, vm = v.copy(), v.copy()
            vp[j, c] += eps
            vm[j, c] -= eps
            fd = (loss(vp) - loss(vm)) / (2 * eps)
            assert abs(fd - grad[j, c]) <= 2e-6 * max(1.0, abs(fd)), (j, c, fd, grad[j, c])


def test_forward_t_matches_traced_t(orc):
    v, f, hit, g, w, h = _setup(orc)
    o = orc.primary_eval64(v, f, hit, g, w, h, np.array([0.0, 1.0, 0.0]))
    tris = scenes.deindex(v.astype(np.float32), f)
    nodes, root, _, _ = orc.build(tris)
    rays = np.zeros((w * h, 8), np.float32)
    for p in range(w * h):
        ro, rd = orc.raygen_shoot(g, (p % w) / w, (p // w) / h)
        rays[p, 0:3], rays[p, 4:7] = ro, rd
    rays[:, 3] = np.float32(3.402823466e38)
    hits, _ = orc.trace(nodes, tris, root, rays)
    m = hit != 0xFFFFFFFF
    assert np.allclose(o[m, 0], hits["t"][m], rtol=1e-5)
    assert np.all(o[:, 1] >= 0) and np.all(o[:, 1] <= 1 + 1e-12)
