"""numpy front end of the CPU oracle (oracle/mlbvh_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libmlbvh_oracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

HIT_DTYPE = np.dtype([("t", "<f4"), ("pad", "<u4"), ("u", "<f4"), ("v", "<f4"), ("ng", "<f4", (3,)),
                      ("tri", "<u4")])
assert HIT_DTYPE.itemsize == 32

FLT_MAX = np.float32(3.402823466e38)
INVALID = 0x7FFFFFFF
LEAF = 0x80000000


def build_lib(force: bool = False) -> str:
    src = os.path.join(HERE, "mlbvh_oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)
    return LIB


class RayGen(ctypes.Structure):
    _fields_ = [("origin", c_float * 3), ("right", c_float * 3), ("up", c_float * 3)]


class V3(ctypes.Structure):
    _fields_ = [("x", c_float), ("y", c_float), ("z", c_float)]


_lib = None


def load(path: str) -> ctypes.CDLL:
    """The oracle library at `path` with its function signatures set."""
    L = ctypes.CDLL(path)
    P = c_void_p
    sig = {
        "orc_clz": ([c_uint32], c_int), "orc_clz64": ([c_uint64], c_int), "orc_ffs": ([c_uint32], c_int),
        "orc_ffs64": ([c_uint64], c_int), "orc_delta32": ([c_uint32, c_uint32], c_int),
        "orc_splat3": ([c_uint32], c_uint64), "orc_compact3": ([c_uint64], c_uint32),
        "orc_encode_morton": ([c_uint32, c_uint32, c_uint32], c_uint64),
        "orc_morton_batch": ([P, P, c_int, c_int], None),
        "orc_decode_batch": ([P, P, c_int], None),
        "orc_f2i_trunc": ([c_float], c_int),
        "orc_scene_aabb": ([P, c_int, P], None),
        "orc_morton_records": ([P, c_int, P, c_uint32, P], None),
        "orc_sort_records": ([P, c_int], None),
        "orc_deltas": ([P, c_int, P], None),
        "orc_build_leaf": ([P, P, P, c_int, P, c_uint32, c_uint32], None),
        "orc_build": ([P, c_int, c_uint32, P, P, P, P, c_int], c_int),
        "orc_validate": ([c_uint32, P, P, c_int], c_int),
        "orc_radix_sort_u32": ([P, P, c_uint32, c_int, c_int], c_int),
        "orc_radix_sort_u64": ([P, P, c_uint32, c_int, c_int], c_int),
        "orc_trace": ([P, P, c_uint32, P, c_int, c_int, c_int, P, P, c_int], None),
        "orc_brute_force": ([P, c_int, P, c_int, P], None),
        "orc_hash_pcg": ([c_uint32], c_uint32),
        "orc_pcg32": ([c_uint64, c_uint64, c_int, P], None),
        "orc_pcg32_uniformf": ([c_uint64, c_uint64, c_int, P], None),
        "orc_hemisphere_cos": ([c_float, c_float, P], None),
        "orc_onb": ([P, P], None),
        "orc_cosine_dir": ([P, c_float, c_float, P], None),
        "orc_sincos": ([c_float, POINTER(c_float), POINTER(c_float)], None),
        "orc_log2_pos": ([c_float], c_float),
        "orc_exp2_fin": ([c_float], c_float),
        "orc_pow_pos": ([c_float, c_float], c_float),
        "orc_sincos_batch": ([P, c_int, P, P], None),
        "orc_points_morton": ([P, c_int, P, P], None),
        "orc_pow_batch": ([P, c_int, c_float, P], None),
        "orc_sobol_2d": ([POINTER(c_float), POINTER(c_float), c_uint32, c_uint32, c_uint32, c_uint32], None),
        "orc_raygen_lookat": ([P, V3, V3, V3, c_float, c_int, c_int], None),
        "orc_raygen_shoot": ([P, P, P, c_float, c_float], None),
        "orc_pack_rgba": ([c_float, c_float, c_float, c_float], c_uint32),
        "orc_render_normal": ([P, c_int, c_int, P, c_uint32, P, P, c_int], None),
        "orc_render_ao": ([P, P, c_int, c_int, P, c_uint32, P, P, c_int, c_int, c_float, c_uint32, c_int,
                           c_int, P, c_int, c_int], None),
        "orc_render_pt": ([P, c_int, c_int, P, c_uint32, P, P, P, c_int, c_int, c_int, c_int], None),
        "orc_any_hit_stack_batch": ([P, P, c_uint32, P, c_int, c_int, P, P, c_int], None),
        "orc_closest_hit_stack_batch": ([P, P, c_uint32, P, c_int, c_int, c_int, P, P, c_int], None),
        "orc_pack": ([P, P, c_int], None),
        "orc_primary_eval64": ([P, P, P, c_int, c_int, P, P, P], None),
        "orc_primary_grad64": ([P, P, P, c_int, c_int, P, P, P, P], None),
        "orc_now": ([], c_double),
        "orc_max_threads": ([], c_int),
    }
    for name, (a, r) in sig.items():
        f = getattr(L, name)
        f.argtypes = a
        f.restype = r
    return L


def lib():
    global _lib
    if _lib is None:
        build_lib()
        _lib = load(LIB)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def morton_batch(xyz, naive=False):
    xyz = np.ascontiguousarray(xyz, np.uint32)
    out = np.zeros(xyz.shape[0], np.uint64)
    lib().orc_morton_batch(_p(xyz), _p(out), xyz.shape[0], 1 if naive else 0)
    return out


def decode_batch(codes):
    codes = np.ascontiguousarray(codes, np.uint64)
    xyz = np.zeros((codes.shape[0], 3), np.uint32)
    lib().orc_decode_batch(_p(codes), _p(xyz), codes.shape[0])
    return xyz


def scene_aabb(tris):
    tris = _f32(tris)
    out = np.zeros(6, np.float32)
    lib().orc_scene_aabb(_p(tris), tris.shape[0], _p(out))
    return out


def morton_records(tris, opt=0, scene=None):
    tris = _f32(tris)
    n = tris.shape[0]
    if scene is None:
        scene = scene_aabb(tris)
    scene = _f32(scene)
    rec = np.zeros(n, np.uint64)
    lib().orc_morton_records(_p(tris), n, _p(scene), opt, _p(rec))
    return rec


def build(tris, opt=0, parallel=True, nthreads=0):
    """Returns (nodes uint32 (n-1,16), root, sorted u64 (n,), deltas u8 (n-1,))."""
    tris = _f32(tris)
    n = tris.shape[0]
    nodes = np.zeros((max(n - 1, 0), 16), np.uint32)
    root = np.zeros(1, np.uint32)
    rec = np.zeros(n, np.uint64)
    deltas = np.zeros(max(n, 1), np.uint8)
    o = opt | (2 if parallel else 0)
    rc = lib().orc_build(_p(tris), n, o, _p(nodes), _p(root), _p(rec), _p(deltas), nthreads)
    if rc != 0:
        raise ValueError(f"oracle build rejected n={n}")
    return nodes, int(root[0]), rec, deltas[: n - 1]


def build_from_deltas(deltas, tri_index=None, tris=None):
    """Run the agglomerative climb serially on given deltas (the toy example of
    main.cpp:60-109 passes no triangles)."""
    deltas = np.ascontiguousarray(deltas, np.uint8)
    n = deltas.shape[0] + 1
    nodes = np.zeros((n - 1, 16), np.uint32)
    nodes[:, 15] = 0xFFFFFFFF
    root = np.zeros(1, np.uint32)
    idx = np.arange(n, dtype=np.uint32) if tri_index is None else np.asarray(tri_index, np.uint32)
    t = None if tris is None else _f32(tris)
    for i in range(n):
        lib().orc_build_leaf(_p(root), _p(nodes), _p(t), n, _p(deltas), i, int(idx[i]))
    return nodes, int(root[0])


def validate(root, nodes, deltas, max_delta=255):
    nodes = np.ascontiguousarray(nodes, np.uint32)
    deltas = np.ascontiguousarray(deltas, np.uint8)
    return bool(lib().orc_validate(root, _p(nodes), _p(deltas), max_delta))


def radix_sort(keys, values=None, start_bit=0, end_bit=None):
    keys = np.array(keys, copy=True)
    if end_bit is None:
        end_bit = keys.dtype.itemsize * 8
    vals = None if values is None else np.array(values, dtype=np.uint32, copy=True)
    fn = lib().orc_radix_sort_u64 if keys.dtype.itemsize == 8 else lib().orc_radix_sort_u32
    passes = fn(_p(keys), _p(vals), keys.shape[0], start_bit, end_bit)
    return keys, vals, passes


def trace(nodes, tris, root, rays, query=0, method=0, nthreads=0):
    tris = _f32(tris)
    rays = _f32(rays)
    nodes = np.ascontiguousarray(nodes, np.uint32)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    cnt = np.zeros(2, np.uint64)
    lib().orc_trace(_p(nodes), _p(tris), root, _p(rays), rays.shape[0], query, method, _p(hits), _p(cnt), nthreads)
    return hits, cnt


def brute_force(tris, rays):
    tris = _f32(tris)
    rays = _f32(rays)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    lib().orc_brute_force(_p(tris), tris.shape[0], _p(rays), rays.shape[0], _p(hits))
    return hits


def sobol_2d(index, p0, p1, p2):
    x, y = c_float(), c_float()
    lib().orc_sobol_2d(ctypes.byref(x), ctypes.byref(y), index, p0, p1, p2)
    return np.float32(x.value), np.float32(y.value)


def onb(n):
    """(xaxis, yaxis) of the Duff et al. 2017 basis for the unit normal n."""
    out = np.zeros(6, np.float32)
    lib().orc_onb(_p(_f32(n)), _p(out))
    return out[:3], out[3:]


def cosine_dir(n, xi0, xi1):
    out = np.zeros(3, np.float32)
    lib().orc_cosine_dir(_p(_f32(n)), xi0, xi1, _p(out))
    return out


def sincos(x):
    s, c = c_float(), c_float()
    lib().orc_sincos(x, ctypes.byref(s), ctypes.byref(c))
    return np.float32(s.value), np.float32(c.value)


def points_morton(pts):
    """(box (6,), 63-bit codes) of AABB::encodeMortonCode over the points' box."""
    pts = _f32(pts).reshape(-1, 3)
    box = np.zeros(6, np.float32)
    codes = np.zeros(pts.shape[0], np.uint64)
    lib().orc_points_morton(_p(pts), pts.shape[0], _p(box), _p(codes))
    return box, codes


def sincos_batch(x):
    x = _f32(x)
    s, c = np.zeros_like(x), np.zeros_like(x)
    lib().orc_sincos_batch(_p(x), x.size, _p(s), _p(c))
    return s, c


def pow_batch(x, y):
    x = _f32(x)
    out = np.zeros_like(x)
    lib().orc_pow_batch(_p(x), x.size, y, _p(out))
    return out


def raygen_lookat(eye, center, up, fovy, w, h):
    g = RayGen()
    lib().orc_raygen_lookat(ctypes.byref(g), V3(*eye), V3(*center), V3(*up), fovy, w, h)
    return g


def raygen_shoot(g, u, v):
    ro, rd = V3(), V3()
    lib().orc_raygen_shoot(ctypes.byref(g), ctypes.byref(ro), ctypes.byref(rd), u, v)
    return np.array([ro.x, ro.y, ro.z], np.float32), np.array([rd.x, rd.y, rd.z], np.float32)


def render_normal(g, w, h, root, nodes, tris, nthreads=0):
    px = np.zeros(w * h, np.uint32)
    lib().orc_render_normal(_p(px), w, h, ctypes.byref(g), root, _p(np.ascontiguousarray(nodes, np.uint32)),
                            _p(_f32(tris)), nthreads)
    return px


def render_ao(g, w, h, root, nodes, tris, use_sobol=1, n_sample=16, radius=1.0, seed=12345, row0=0, row1=None,
              nthreads=0, early_exit=False, ao_stack=False, primary_stack=False):
    """walk: the reference stack-free walk; early_exit: with the pending-count
    exit; ao_stack: early-exit primary ray + stack any-hit AO rays;
    primary_stack: stack closest-hit primary ray + stack any-hit AO rays (the
    GPU kernel's walks, so its counters match)."""
    px = np.zeros(w * h, np.uint32)
    ao = np.full(w * h, np.nan, np.float32)
    cnt = np.zeros(3, np.uint64)
    walk = 3 if primary_stack else (2 if ao_stack else (1 if early_exit else 0))
    lib().orc_render_ao(_p(px), _p(ao), w, h, ctypes.byref(g), root, _p(np.ascontiguousarray(nodes, np.uint32)),
                        _p(_f32(tris)), use_sobol, n_sample, radius, seed & 0xFFFFFFFF, row0,
                        h if row1 is None else row1, _p(cnt), nthreads, walk)
    return px, ao, cnt


def closest_hit_stack(nodes, tris, root, rays, query=0, cap=16, nthreads=0):
    """Hit records and (node fetches, triangle tests) of the stack walk with
    pop re-test, stack of `cap` entries (overflow restarts stack-free)."""
    rays = _f32(rays)
    hits = np.zeros(rays.shape[0], HIT_DTYPE)
    cnt = np.zeros(2, np.uint64)
    lib().orc_closest_hit_stack_batch(_p(np.ascontiguousarray(nodes, np.uint32)), _p(_f32(tris)), root, _p(rays),
                                      rays.shape[0], query, cap, _p(hits), _p(cnt), nthreads)
    return hits, cnt


def any_hit_stack(nodes, tris, root, rays, cap=16, nthreads=0):
    """Occluded flags (bool, one per ray) and (node fetches, triangle tests) of
    the stack any-hit walk with a stack of `cap` entries."""
    rays = _f32(rays)
    n = rays.shape[0]
    occ = np.zeros(n, np.uint8)
    cnt = np.zeros(2, np.uint64)
    lib().orc_any_hit_stack_batch(_p(np.ascontiguousarray(nodes, np.uint32)), _p(_f32(tris)), root, _p(rays), n, cap,
                                  _p(occ), _p(cnt), nthreads)
    return occ.astype(bool), cnt


def render_pt(acc, g, w, h, root, nodes, tris, attribs, use_sobol=1, max_spp=1024, max_depth=1024, nthreads=0):
    acc = np.ascontiguousarray(acc, np.float32)
    lib().orc_render_pt(_p(acc), w, h, ctypes.byref(g), root, _p(np.ascontiguousarray(nodes, np.uint32)),
                        _p(_f32(tris)), _p(_f32(attribs)), use_sobol, max_spp, max_depth, nthreads)
    return acc


def pack(acc):
    acc = np.ascontiguousarray(acc, np.float32)
    n = acc.size // 4
    px = np.zeros(n, np.uint32)
    lib().orc_pack(_p(px), _p(acc), n)
    return px


def primary_eval64(verts, idx, hit_tri, g, w, h, light):
    """fp64 (t, shade) per pixel with the hit triangles held fixed."""
    v = np.ascontiguousarray(verts, np.float64)
    out = np.zeros((w * h, 2), np.float64)
    lib().orc_primary_eval64(_p(v), _p(np.ascontiguousarray(idx, np.uint32)),
                             _p(np.ascontiguousarray(hit_tri, np.uint32)), w, h, ctypes.byref(g),
                             _p(np.ascontiguousarray(light, np.float64)), _p(out))
    return out


def primary_grad64(verts, idx, hit_tri, g, w, h, light, gout):
    v = np.ascontiguousarray(verts, np.float64)
    grad = np.zeros(v.shape, np.float64)
    lib().orc_primary_grad64(_p(v), _p(np.ascontiguousarray(idx, np.uint32)),
                             _p(np.ascontiguousarray(hit_tri, np.uint32)), w, h, ctypes.byref(g),
                             _p(np.ascontiguousarray(light, np.float64)),
                             _p(np.ascontiguousarray(gout, np.float64)), _p(grad))
    return grad


def now():
    return lib().orc_now()


def max_threads():
    return lib().orc_max_threads()
