"""The synthetic workloads of the benchmark, each runnable on its own so a
profiler can wrap exactly one of them:

    python -m minimumlbvh_amd.workloads ao       [--reps K]   # 1080p AO frames, level-9 sphere
    python -m minimumlbvh_amd.workloads build    [--reps K]   # BVH build of the level-9 sphere
    python -m minimumlbvh_amd.workloads sort     [--reps K]   # 2^25 uniform u64 keys, 64 bits
    python -m minimumlbvh_amd.workloads batched  [--reps K]   # 4096 meshes, one build + trace
    python -m minimumlbvh_amd.workloads diff     [--reps K]   # level-7 sphere, 512^2, fwd+bwd steps
    python -m minimumlbvh_amd.workloads walks    [--reps K]   # escape-link / indexed trace, indexed / 63-bit build
    python -m minimumlbvh_amd.workloads pt       [--reps K]   # 21M-tri sphere field: build + 1080p 8 spp PT
    python -m minimumlbvh_amd.workloads copy     [--reps K]   # 1 GiB device copy (counter calibration)
    python -m minimumlbvh_amd.workloads gather   [--reps K]   # random 64-B node gathers (traversal roof)

Scenes:
  * ao / build: procedural icosphere of subdivision level 9 (5,242,880
    triangles) with radial displacement 0.05 x (4 seeded sine octaves),
    de-indexed to 36-byte Triangle records; camera CAMERA_SPHERE, 1920x1080,
    1 primary + 16 AO rays per hit pixel, radius 0.2, Sobol seed 0.
  * batched: 4096 meshes of independent random triangles (centroids uniform
    in [0,1]^3, vertices at N(0, 0.01^2) offsets), per-mesh triangle counts
    log-uniform in [64, 262144] and ray counts log-uniform in [256, 65536]
    (seeded, scenes.batch_sizes); rays start on the sphere of radius 2 about
    (0.5,0.5,0.5) and aim at it.  Generated on the device from a seeded torch
    generator (about 1.3e8 triangles, 4.7 GB).
  * diff: level-7 displaced icosphere (327,680 triangles), 512x512, depth +
    n.l shading, L2 loss against the mesh with an extra seeded displacement
    of amplitude 0.02 over the pixels that hit in both images, Adam lr 1e-3,
    refit every step, rebuild every 20.
"""
from __future__ import annotations

import argparse
import math
import sys
import time

import numpy as np
import torch

from . import autotune, lbvh, scenes

AO_W, AO_H, AO_SPP, AO_RADIUS, AO_SEED = 1920, 1080, 16, 0.2, 0
FLT_MAX = 3.402823466e38
LIGHT = (0.3 / math.sqrt(0.3**2 + 0.8**2 + 0.52**2), 0.8 / math.sqrt(0.3**2 + 0.8**2 + 0.52**2),
         -0.52 / math.sqrt(0.3**2 + 0.8**2 + 0.52**2))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class AOScene:
    """Level-`level` displaced icosphere with its GPU BVH resident in HBM."""

    def __init__(self, level: int = 9, device: str = "cuda:0"):
        self.verts, self.faces = scenes.displaced_icosphere(level, seed=0)
        self.tris_np = scenes.deindex(self.verts, self.faces)
        self.n = self.tris_np.shape[0]
        self.tris = torch.from_numpy(self.tris_np).to(device)
        self.sorter = lbvh.OnesweepSort()
        self.builder = lbvh.BVHGPUBuilder()
        self.eye, self.ctr, self.up, fov = scenes.CAMERA_SPHERE
        self.fovy = math.radians(fov)
        self.cam = lbvh.RayGenerator().lookat(self.eye, self.ctr, self.up, self.fovy, AO_W, AO_H)
        self.px = torch.zeros(AO_W * AO_H, dtype=torch.int32, device=device)

    def build(self, stream=None):
        self.builder.build(self.tris, self.n, 0, self.sorter, stream)

    def frame(self, counters=None, stream=None):
        lbvh.render_ao(self.px, AO_W, AO_H, self.cam, self.builder.m_rootNode, self.builder.m_internals, self.tris,
                       1, AO_SPP, AO_RADIUS, AO_SEED, counters=counters, stream=stream)


def random_u64(n: int, seed: int = 0, device: str = "cuda:0") -> torch.Tensor:
    g = torch.Generator(device=device).manual_seed(seed)
    return torch.randint(-(2**63), 2**63 - 1, (n,), dtype=torch.int64, device=device, generator=g)


class BatchedScene:
    """The many-mesh workload, generated on the device."""

    def __init__(self, n_meshes: int = 4096, seed: int = 0, device: str = "cuda:0",
                 tri_range=(64, 262144), ray_range=(256, 65536)):
        tc, rc = scenes.batch_sizes(n_meshes, seed=seed, tri_range=tri_range, ray_range=ray_range)
        self.toff = np.concatenate([[0], np.cumsum(tc)]).astype(np.uint32)
        self.roff = np.concatenate([[0], np.cumsum(rc)]).astype(np.uint32)
        T, R = int(self.toff[-1]), int(self.roff[-1])
        self.n_tris, self.n_rays, self.n_meshes = T, R, n_meshes
        g = torch.Generator(device=device).manual_seed(seed)
        c = torch.rand((T, 1, 3), generator=g, device=device)
        self.tris = (c + 0.01 * torch.randn((T, 3, 3), generator=g, device=device)).contiguous()
        del c
        d = torch.randn((R, 3), generator=g, device=device, dtype=torch.float64)
        d /= d.norm(dim=1, keepdim=True)
        o = 0.5 + 2.0 * d
        rays = torch.zeros((R, 8), dtype=torch.float32, device=device)
        rays[:, 0:3] = o.float()
        rays[:, 3] = FLT_MAX
        rays[:, 4:7] = (0.5 - o).float()
        mesh = torch.repeat_interleave(torch.arange(n_meshes, dtype=torch.int32, device=device),
                                       torch.from_numpy(rc).to(device))
        rays[:, 7] = mesh.view(torch.float32)
        self.rays = rays.contiguous()
        self.bvh = lbvh.BatchedBVH(self.toff)
        self.hits = torch.empty((R, 8), dtype=torch.int32, device=device)

    def build_trace(self, stream=None):
        self.bvh.build(self.tris, 0, stream)
        self.bvh.trace(self.tris, self.rays, hits=self.hits, stream=stream)

    def per_mesh_loop(self, hits=None, stream=None):
        return self.bvh.build_trace_per_mesh(self.tris, self.rays, self.roff, hits=hits, stream=stream)


class DiffScene:
    """Level-7 sphere fitted to a displaced target through the adjoint."""

    def __init__(self, level: int = 7, w: int = 512, h: int = 512, rebuild_every: int = 20,
                 device: str = "cuda:0"):
        from . import diff

        v, f = scenes.displaced_icosphere(level, seed=0)
        dirs = v / np.linalg.norm(v, axis=1, keepdims=True)
        v_tgt = (v * scenes.displacement(dirs, 99, 0.02)[:, None].astype(np.float32)).astype(np.float32)
        eye, ctr, up, fov = scenes.CAMERA_SPHERE
        cam = lbvh.RayGenerator().lookat(eye, ctr, up, math.radians(fov), w, h)
        faces = torch.from_numpy(f.astype(np.int32)).to(device)
        self.renderer = diff.PrimaryRenderer(faces, w, h, cam, LIGHT, rebuild_every=rebuild_every)
        self.target = self.renderer.forward_raw(torch.from_numpy(v_tgt).to(device)).detach().clone()
        self.renderer.calls = 0
        self.v0 = torch.from_numpy(v).to(device)
        self.vt = self.v0.clone().requires_grad_(True)
        self.opt = torch.optim.Adam([self.vt], lr=1e-3)
        self.depth_w = 0.5

    def loss_of(self, out):
        # pixels that hit in both images: with the hit triangle held fixed the
        # gradient has no term for a silhouette that moves, so a pixel that
        # hits in one image only would pull its triangle along the ray
        both = ((out[..., 0] > 0) & (self.target[..., 0] > 0)).to(out.dtype)
        return self.depth_w * (both * (out[..., 0] - self.target[..., 0]) ** 2).sum() + \
            (both * (out[..., 1] - self.target[..., 1]) ** 2).sum()

    def forward(self):
        with torch.no_grad():
            return self.loss_of(self.renderer.forward_raw(self.vt.detach()))

    def step(self):
        self.opt.zero_grad(set_to_none=False)
        loss = self.loss_of(self.renderer.render(self.vt))
        loss.backward()
        self.opt.step()
        return loss


def events_ms(fn, reps: int, warmup: int = 1, stream=None):
    """Average ms per call of fn over reps, from HIP events on `stream`."""
    stream = stream or torch.cuda.current_stream()
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record(stream)
    for _ in range(reps):
        fn()
    e.record(stream)
    torch.cuda.synchronize()
    return s.elapsed_time(e) / reps


_calib = None


def calib_lib():
    """The roofline probes of csrc/calib/probes.hip (libmlbvh_calib.so)."""
    global _calib
    if _calib is None:
        import ctypes
        import os

        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libmlbvh_calib.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is missing: run `python -m minimumlbvh_amd._build` first")
        L = ctypes.CDLL(path)
        L.mlbvh_calib_gather.argtypes = [ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
        L.mlbvh_calib_gather.restype = ctypes.c_int
        _calib = L
    return _calib


# Gather probes: (name, table bytes, dependent, blocks per CU).  2 MiB fits
# in every XCD's 4 MiB L2; 256 MiB is the size of the MI355X's Infinity
# Cache (MALL); 4 GiB exceeds both (HBM).  The dependent probe at 5 blocks of
# 256 lanes per CU is the AO kernel's residency (20 waves per CU).
GATHER_PROBES = [("l2_independent", 2 << 20, 0, 8), ("l2_dependent_ao_occupancy", 2 << 20, 1, 5),
                 ("mall_independent", 256 << 20, 0, 8), ("hbm_independent", 4 << 30, 0, 8),
                 ("l2_quad_cooperative", 2 << 20, 2, 8), ("mall_quad_cooperative", 256 << 20, 2, 8)]


def run_gather(reps: int, out: dict, iters: int = 256, probes=None):
    """Random 64-B node gathers (four 16-B loads each, like the traversal's
    node fetch) from tables of several sizes: GB/s of 64-B node fetches."""
    import ctypes

    L = calib_lib()
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    stream = torch.cuda.current_stream()
    for name, nbytes, dep, bpc in probes or GATHER_PROBES:
        nodes = nbytes // 64
        table = torch.randint(0, 2**31 - 1, (nbytes // 4,), dtype=torch.int32, device="cuda:0")
        grid = cus * bpc
        sink = torch.empty(grid * 256, dtype=torch.int32, device="cuda:0")

        def run():
            rc = L.mlbvh_calib_gather(ctypes.c_void_p(table.data_ptr()), nodes, iters, dep,
                                      ctypes.c_void_p(sink.data_ptr()), grid, ctypes.c_void_p(stream.cuda_stream))
            if rc != 0:
                raise RuntimeError(f"mlbvh_calib_gather failed ({rc})")

        ms = events_ms(run, reps, warmup=2, stream=stream)
        fetched = grid * 256 * iters * 64
        out[name] = {"table_bytes": nbytes, "dependent": bool(dep), "lanes": grid * 256, "fetches_per_lane": iters,
                     "ms": round(ms, 4), "gbs": round(fetched / (ms / 1e3) / 1e9, 1)}
        del table, sink
    return out


def run_batched(reps: int, out: dict, graph: bool = True):
    """C3: one segmented build + persistent trace of all meshes vs the plain
    per-mesh launch loop; the batched pipeline also captured in a hipGraph."""
    log("[batched] generating 4096 meshes on the device")
    sc = BatchedScene()
    out.update(meshes=sc.n_meshes, triangles=sc.n_tris, rays=sc.n_rays)
    sc.build_trace()
    torch.cuda.synchronize()
    best, res = autotune.tune_trace(sc.bvh, sc.tris, sc.rays)
    out["trace_shape"] = {"blocks_per_cu": best[0], "chunk": best[1],
                          "ms_by_shape": {f"{k[0]}x{k[1]}": round(v, 3) for k, v in res.items()}}
    batched_ms = events_ms(sc.build_trace, reps)
    out["batched_ms"] = round(batched_ms, 3)
    ref = sc.hits.clone()
    if graph:
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            sc.build_trace(stream=s)  # sizes every buffer before capture
        torch.cuda.current_stream().wait_stream(s)
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            sc.build_trace()
        graph_ms = events_ms(g.replay, reps)
        torch.cuda.synchronize()
        out["graph_ms"] = round(graph_ms, 3)
        out["graph_hits_equal"] = bool(torch.equal(sc.hits, ref))
    loop_hits = torch.empty_like(sc.hits)
    sc.per_mesh_loop(loop_hits)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    loop_reps = max(1, min(reps, 2))
    for _ in range(loop_reps):
        sc.per_mesh_loop(loop_hits)
    torch.cuda.synchronize()
    loop_ms = (time.perf_counter() - t0) * 1e3 / loop_reps
    out["per_mesh_loop_ms"] = round(loop_ms, 3)
    out["loop_hits_equal"] = bool(torch.equal(loop_hits, ref))
    out["speedup"] = round(loop_ms / min(batched_ms, out.get("graph_ms", batched_ms)), 2)
    return sc


def sphere_rays(n: int, seed: int = 0, tmax: float = FLT_MAX, device: str = "cuda:0") -> torch.Tensor:
    """n rays from the sphere of radius 3 about the origin towards points
    uniform in the unit ball (the level-9 sphere has radius about 1)."""
    g = torch.Generator(device=device).manual_seed(seed)
    d = torch.randn((n, 3), generator=g, device=device, dtype=torch.float64)
    o = 3.0 * d / d.norm(dim=1, keepdim=True)
    t = torch.randn((n, 3), generator=g, device=device, dtype=torch.float64)
    t = t / t.norm(dim=1, keepdim=True) * torch.rand((n, 1), generator=g, device=device, dtype=torch.float64) ** (1 / 3)
    rays = torch.zeros((n, 8), dtype=torch.float32, device=device)
    rays[:, 0:3] = o.float()
    rays[:, 3] = tmax
    rays[:, 4:7] = (t - o).float()
    return rays.contiguous()


def run_walks(reps: int, out: dict, sc: "AOScene | None" = None):
    """The alternatives beside the headline path, on the level-9 sphere: the
    stack walk, the stack-free walk, the escape-link walk and the indexed-mesh
    trace on 4M closest-hit rays, and the build from the index buffer and
    with 63-bit keys."""
    sc = sc or AOScene()
    sc.build()
    n = 1 << 22
    rays = sphere_rays(n)
    vt = torch.from_numpy(sc.verts).to(rays.device)
    ft = torch.from_numpy(sc.faces.astype(np.int32)).to(rays.device)
    walks = {}
    hits = {}
    for name, fn in (("stack", lambda c=None: lbvh.trace(sc.builder, sc.tris, rays, counters=c)),
                     ("stackfree", lambda c=None: lbvh.trace(sc.builder, sc.tris, rays, counters=c,
                                                             walk="stackfree")),
                     ("escape_link", lambda c=None: lbvh.trace(sc.builder, sc.tris, rays, counters=c,
                                                               walk="escape_link")),
                     ("indexed", lambda c=None: lbvh.trace_indexed(sc.builder, vt, ft, rays, counters=c))):
        cnt = torch.zeros(2, dtype=torch.int64, device=rays.device)
        hits[name] = fn(cnt)
        ms = events_ms(fn, reps)
        c = cnt.cpu().numpy().view(np.uint64)
        walks[name] = {"ms": round(ms, 4), "mrays_per_s": round(n / ms / 1e3, 1), "node_visits": int(c[0]),
                       "triangle_tests": int(c[1])}
    torch.cuda.synchronize()
    out["rays"] = n
    out["trace"] = walks
    out["stackfree_hits_equal"] = bool(torch.equal(hits["stack"], hits["stackfree"]))
    out["escape_t_equal"] = bool(torch.equal(hits["stack"][:, 0], hits["escape_link"][:, 0]))
    out["indexed_hits_equal"] = bool(torch.equal(hits["stack"], hits["indexed"]))
    ref = sc.builder.m_internals.clone()
    b = lbvh.BVHGPUBuilder()
    b.build_indexed(vt, ft)
    torch.cuda.synchronize()
    out["indexed_build_nodes_equal"] = bool(torch.equal(b.m_internals, ref))
    out["build_ms"] = round(events_ms(sc.build, reps), 4)
    out["indexed_build_ms"] = round(events_ms(lambda: b.build_indexed(vt, ft), reps), 4)
    out["morton64_build_ms"] = round(events_ms(lambda: b.build(sc.tris, sc.n, lbvh.BUILD_OPTION_MORTON64), reps), 4)
    return sc


def run_pt(reps: int, out: dict, spp: int = 8, max_depth: int = 4):
    """The path-tracing scene: 64 displaced level-7 spheres + ground + light
    (20,971,524 triangles) built on the GPU, then `spp` samples per pixel at
    1920x1080, max depth 4, Sobol streams; time per build and per frame of
    `spp` samples."""
    tris_np, attribs_np = scenes.pt_scene()
    dev = "cuda:0"
    tris = torch.from_numpy(tris_np).to(dev)
    attribs = torch.from_numpy(attribs_np).to(dev)
    n = tris_np.shape[0]
    del tris_np, attribs_np
    sorter, builder = lbvh.OnesweepSort(), lbvh.BVHGPUBuilder()
    build = lambda: builder.build(tris, n, 0, sorter)  # noqa: E731
    build()
    build_ms = events_ms(build, reps)
    eye, ctr, up, fov = scenes.CAMERA_PT
    w, h = AO_W, AO_H
    cam = lbvh.RayGenerator().lookat(eye, ctr, up, math.radians(fov), w, h)
    acc = torch.zeros((w * h, 4), dtype=torch.float32, device=dev)

    def frame():
        acc.zero_()
        for _ in range(spp):
            lbvh.render_pt(acc, w, h, cam, builder.m_rootNode, builder.m_internals, tris, attribs, 1, 1024, max_depth)

    frame_ms = events_ms(frame, reps)
    out.update(triangles=n, width=w, height=h, spp=spp, max_depth=max_depth, build_ms=round(build_ms, 4),
               build_mtris_per_s=round(n / build_ms / 1e3, 1), frame_ms=round(frame_ms, 3),
               mpaths_per_s=round(w * h * spp / frame_ms / 1e3, 1),
               mean_radiance=float(acc[:, :3].mean()))


def run_diff(steps: int, out: dict):
    """C4: Adam steps (fwd + bwd + update) against forward-only time.  One
    untimed step first: the first launch of each backward, refit and rebuild
    kernel loads its code, which is not per-step work."""
    sc = DiffScene()
    fwd_ms = events_ms(sc.forward, 20, 2)
    l0 = float(sc.forward())
    sc.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev_s, ev_e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev_s.record()
    for _ in range(steps):
        loss = sc.step()
    ev_e.record()
    torch.cuda.synchronize()
    step_ms = ev_s.elapsed_time(ev_e) / steps
    wall_ms = (time.perf_counter() - t0) * 1e3 / steps
    l1 = float(sc.forward())
    out.update(triangles=sc.renderer.n_tris, pixels=sc.renderer.width * sc.renderer.height, adam_steps=steps,
               forward_ms=round(fwd_ms, 4), fwd_bwd_step_ms=round(step_ms, 4), step_wall_ms=round(wall_ms, 4),
               ratio=round(step_ms / fwd_ms, 3), loss_first=l0, loss_last=l1, last_step_loss=float(loss))
    return sc


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("workload", choices=["ao", "build", "sort", "batched", "diff", "walks", "pt", "copy", "gather"])
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--probe", default=None, help="gather: run only this probe (names in GATHER_PROBES)")
    args = ap.parse_args(argv)
    res = {"workload": args.workload}
    if args.workload in ("ao", "build"):
        sc = AOScene()
        sc.build()
        if args.workload == "build":
            res["ms"] = events_ms(sc.build, args.reps)
        else:
            res["ms"] = events_ms(sc.frame, args.reps)
    elif args.workload == "sort":
        n = 1 << 25
        src = random_u64(n)
        keys, tmp = torch.empty_like(src), torch.empty_like(src)
        sorter = lbvh.OnesweepSort()

        def run():
            keys.copy_(src)
            sorter.sort(keys, None, tmp, None, n, 0, 64)

        res["ms_incl_copy"] = events_ms(run, args.reps)
    elif args.workload == "batched":
        run_batched(args.reps, res, graph=False)
    elif args.workload == "diff":
        run_diff(max(args.reps, 1), res)
    elif args.workload == "walks":
        run_walks(args.reps, res)
    elif args.workload == "pt":
        run_pt(args.reps, res)
    elif args.workload == "gather":
        pick = [p for p in GATHER_PROBES if args.probe in (None, p[0])]
        run_gather(args.reps, res, probes=pick)
    else:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        res["bytes_read"] = res["bytes_written"] = a.numel() * 4
        res["ms"] = events_ms(lambda: b.copy_(a), args.reps)
    torch.cuda.synchronize()
    print(res, flush=True)


if __name__ == "__main__":
    main()
