"""Seeded synthetic scenes for the workloads (numpy, host side).

  random_triangles  independent triangles: centroids U[0,1]^3, each vertex
                    = centroid + N(0, sigma^2) per axis (65,536 tris with
                    sigma 0.01 is the small plain-method workload)
  displaced_icosphere  icosahedron subdivided `level` times (level 9 =
                    5,242,880 tris), projected to the unit sphere, radius
                    scaled by 1 + amp * sum of 4 seeded sine octaves
  pt_scene          64 displaced level-7 spheres on a ground quad with one
                    emissive quad (path-tracing workload)

Generators return float32 arrays; ``deindex`` turns an indexed mesh into the
36-byte Triangle records the builder consumes.
"""
from __future__ import annotations

import math

import numpy as np


def random_triangles(n: int, seed: int = 0, sigma: float = 0.01) -> np.ndarray:
    rng = np.random.default_rng(seed)
    c = rng.random((n, 1, 3))
    v = c + rng.normal(0.0, sigma, size=(n, 3, 3))
    return np.ascontiguousarray(v.astype(np.float32))


def icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(level: int):
    """Unit icosphere: 20*4^level faces, 10*4^level+2 vertices (float64, int64).
    The four children of a face are stored consecutively, so face order stays
    spatially coherent."""
    v, f = icosahedron()
    for _ in range(level):
        nv = v.shape[0]
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]], axis=0)
        lo = np.minimum(e[:, 0], e[:, 1])
        hi = np.maximum(e[:, 0], e[:, 1])
        key = lo * nv + hi
        uniq, inv = np.unique(key, return_inverse=True)
        ulo, uhi = uniq // nv, uniq % nv
        mid = v[ulo] + v[uhi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid], axis=0)
        nf = f.shape[0]
        m01 = inv[0:nf] + nv
        m12 = inv[nf:2 * nf] + nv
        m20 = inv[2 * nf:3 * nf] + nv
        f = np.stack(
            [np.stack([f[:, 0], m01, m20], 1), np.stack([f[:, 1], m12, m01], 1), np.stack([f[:, 2], m20, m12], 1),
             np.stack([m01, m12, m20], 1)], axis=1).reshape(-1, 3)
    return v, f


def displacement(dirs: np.ndarray, seed: int, amp: float) -> np.ndarray:
    """1 + amp * sum_k 0.5^k sin(freq_k * <axis_k, dir> + phase_k), k = 0..3."""
    rng = np.random.default_rng(seed)
    r = np.ones(dirs.shape[0], dtype=np.float64)
    for k in range(4):
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        freq = 3.0 * (2.0 ** k) * rng.uniform(0.75, 1.25)
        phase = rng.uniform(0.0, 2.0 * math.pi)
        r += amp * (0.5 ** k) * np.sin(freq * (dirs @ axis) + phase)
    return r


def displaced_icosphere(level: int, seed: int = 0, amp: float = 0.05):
    """Returns (vertices float32 (V,3), faces uint32 (F,3))."""
    v, f = icosphere(level)
    v = v * displacement(v, seed, amp)[:, None]
    return np.ascontiguousarray(v.astype(np.float32)), np.ascontiguousarray(f.astype(np.uint32))


def deindex(verts: np.ndarray, faces: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(verts[faces.astype(np.int64)].astype(np.float32))


def quad(p0, p1, p2, p3) -> np.ndarray:
    return np.array([[p0, p1, p2], [p0, p2, p3]], dtype=np.float32)


def pt_scene(n_spheres: int = 64, level: int = 7, seed: int = 0):
    """Path-tracing scene: spheres centred in [-10,10] x [0,4] x [-10,10]
    (y up), radii U(0.5,2), Lambert albedo U(0.2,0.8), a 40x40 ground quad at
    y=0 (albedo 0.5) and one emissive 6x6 quad light at y=12.
    Returns (triangles (n,3,3) float32, attribs (n,15) float32 TriangleAttrib)."""
    rng = np.random.default_rng(seed)
    base_v, base_f = icosphere(level)
    tris, attrs = [], []
    for s in range(n_spheres):
        c = np.array([rng.uniform(-10, 10), rng.uniform(0, 4), rng.uniform(-10, 10)])
        rad = rng.uniform(0.5, 2.0)
        alb = rng.uniform(0.2, 0.8, size=3)
        v = base_v * displacement(base_v, seed * 1000 + s + 1, 0.05)[:, None] * rad + c
        t = v[base_f].astype(np.float32)
        a = np.zeros((t.shape[0], 15), dtype=np.float32)
        a[:, 9:12] = alb
        tris.append(t)
        attrs.append(a)
    g = quad([-20, 0, -20], [20, 0, -20], [20, 0, 20], [-20, 0, 20])
    ga = np.zeros((2, 15), dtype=np.float32)
    ga[:, 9:12] = 0.5
    light = quad([-3, 12, -3], [-3, 12, 3], [3, 12, 3], [3, 12, -3])
    la = np.zeros((2, 15), dtype=np.float32)
    la[:, 9:12] = 0.0
    la[:, 12:15] = 20.0
    tris += [g, light]
    attrs += [ga, la]
    return np.ascontiguousarray(np.concatenate(tris)), np.ascontiguousarray(np.concatenate(attrs))


def batch_sizes(n_meshes: int, seed: int = 0, tri_range=(64, 262144), ray_range=(256, 65536)):
    """Per-mesh triangle and ray counts, log-uniform in the given ranges
    (4096 meshes with the defaults is the batching workload)."""
    rng = np.random.default_rng(seed)
    lt = np.log(np.array(tri_range, np.float64))
    lr = np.log(np.array(ray_range, np.float64))
    tris = np.round(np.exp(rng.uniform(lt[0], lt[1], n_meshes))).astype(np.int64)
    rays = np.round(np.exp(rng.uniform(lr[0], lr[1], n_meshes))).astype(np.int64)
    return np.clip(tris, 2, None), rays


def batch_scene(tri_counts, ray_counts, seed: int = 0, sigma: float = 0.01):
    """Concatenated independent random-triangle meshes (each like
    random_triangles) and, per mesh, rays from uniform points on the sphere of
    radius 2 about the mesh centre (0.5,0.5,0.5) aimed at the centre.  Ray pad
    word = mesh id.  Returns (tris (T,3,3), tri_offsets (M+1,), rays (R,8),
    ray_offsets (M+1,))."""
    rng = np.random.default_rng(seed)
    tri_counts = np.asarray(tri_counts, np.int64)
    ray_counts = np.asarray(ray_counts, np.int64)
    toff = np.concatenate([[0], np.cumsum(tri_counts)]).astype(np.uint32)
    roff = np.concatenate([[0], np.cumsum(ray_counts)]).astype(np.uint32)
    T, R = int(toff[-1]), int(roff[-1])
    c = rng.random((T, 1, 3), dtype=np.float32)
    tris = (c + rng.normal(0.0, sigma, size=(T, 3, 3)).astype(np.float32)).astype(np.float32)
    d = rng.normal(size=(R, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    centre = np.array([0.5, 0.5, 0.5])
    o = centre + 2.0 * d
    rays = np.zeros((R, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = np.float32(3.402823466e38)
    rays[:, 4:7] = centre - o
    mesh_id = np.repeat(np.arange(len(ray_counts), dtype=np.uint32), ray_counts)
    rays[:, 7] = mesh_id.view(np.float32)
    return np.ascontiguousarray(tris), toff, rays, roff


def furnace_box(albedo: float = 0.5, emission: float = 1.0):
    """The closed cube [-1, 1]^3 as 12 triangles, every face with the same
    reflectance and emission: seen from inside, every path hits a wall at
    every depth, so a path tracer's radiance has a closed form.  Returns
    (tris (12, 3, 3), attribs (12, 15): shading normals, reflectance,
    emissive)."""
    quads = [((-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1)), ((-1, -1, 1), (-1, 1, 1), (1, 1, 1), (1, -1, 1)),
             ((-1, -1, -1), (-1, 1, -1), (-1, 1, 1), (-1, -1, 1)), ((1, -1, -1), (1, -1, 1), (1, 1, 1), (1, 1, -1)),
             ((-1, -1, -1), (-1, -1, 1), (1, -1, 1), (1, -1, -1)), ((-1, 1, -1), (1, 1, -1), (1, 1, 1), (-1, 1, 1))]
    tris = []
    for a, b, c, d in quads:
        tris += [(a, b, c), (a, c, d)]
    tris = np.asarray(tris, np.float32)
    attribs = np.zeros((tris.shape[0], 15), np.float32)
    attribs[:, 9:12] = albedo
    attribs[:, 12:15] = emission
    return tris, attribs


# Camera inside furnace_box, off the centre
CAMERA_FURNACE = ((0.1, 0.2, 0.3), (0.1, 0.2, -1.0), (0.0, 1.0, 0.0), 90.0)
# Cameras of the workloads: (eye, center, up, fovy_degrees)
CAMERA_SMALL = ((0.5, 0.5, -1.5), (0.5, 0.5, -0.5), (0.0, 1.0, 0.0), 60.0)
CAMERA_SPHERE = ((0.0, 0.2, -1.8), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0)
CAMERA_PT = ((0.0, 8.0, -26.0), (0.0, 1.0, 0.0), (0.0, 1.0, 0.0), 60.0)
