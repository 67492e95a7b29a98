#!/usr/bin/env python3
"""Benchmark of the MinimumLBVH GPU path on one MI355X.

Workload (the scale configuration): procedural displaced icosphere, level 9
(5,242,880 triangles, de-indexed to 36-byte Triangle records), BVH built by
the GPU builder, then ambient occlusion at 1920x1080 with 1 primary + 16 AO
rays per hit pixel (ray length 0.2, Owen-scrambled Sobol seed 0).

A step is one AO frame (primary + AO rays) over the resident BVH; `value` is
rays traced per second over the K timed steps.  The same run also times:
  * the full BVH build of that mesh (Mtris/s) and the standalone sort of 2^25
    uniform u64 keys, each against the 8 TB/s HBM roofline;
  * the 4096-mesh batched build + persistent trace, eager and replayed from a
    captured graph, against the plain per-mesh launch loop;
  * on the same mesh, 4M closest-hit rays through the stack-free, the
    escape-link and the indexed-mesh traversal, and the build from the
    index buffer and with 63-bit keys;
  * the path-tracing scene (64 displaced level-7 spheres, 20,971,524
    triangles): GPU build and 1080p frames of 8 spp, max depth 4;
  * 200 Adam steps of the differentiable primary render (level-7 sphere,
    512x512) against its forward-only time;
  * the roof of the AO kernel: random 64-B node gathers from an L2-resident
    table (per-lane and quad-cooperative fetch), measured in the same run;
  * the CPU oracle on this host's cores: the whole AO frame, and the build in
    the reference CPU builder's structure (std::sort, leaf-parallel climb).
Workload set-up lives in minimumlbvh_amd/workloads.py.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batched] [--no-pt] [--no-diff]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AO MRays/s @4M tris 1080p; build Mtris/s; batched & fwd+bwd vs plain"
HBM_PEAK_GBS = 8000.0
BUILD_BYTES_PER_TRI = 306  # scene AABB 36 + keys 44 + 4-pass sort 72 + deltas 9 + init 4 + climb 141
NODE_BYTES, TRI_BYTES = 64, 36


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def timed(fn, steps, warmup):
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in evs:
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1000.0 / steps
    ev_ms = sum(s.elapsed_time(e) for s, e in evs) / steps
    return wall_ms, ev_ms


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--level", type=int, default=9, help="icosphere subdivision level (9 = 5,242,880 tris)")
    ap.add_argument("--sort-log2", type=int, default=25, help="standalone sort size = 2^k u64 keys")
    ap.add_argument("--no-batched", action="store_true", help="skip the 4096-mesh batched leg")
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-render leg")
    ap.add_argument("--no-pt", action="store_true", help="skip the 21M-triangle path-tracing leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); this path has no multi-GPU "
              "component, see DESIGN.md for how it would shard", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from minimumlbvh_amd import autotune, workloads
    from minimumlbvh_amd.workloads import AO_H as H, AO_RADIUS as RADIUS, AO_SEED as SEED, AO_SPP as SPP, \
        AO_W as W

    dev = torch.device("cuda:0")
    log("[bench] autotuning the sort tile")
    best_cfg, cfg_ms = autotune.tune_sort(1 << 24, 64)
    log(f"[bench] sort tile ms by config: {cfg_ms} -> {best_cfg}")
    log(f"[bench] generating icosphere level {args.level}")
    sc = workloads.AOScene(args.level)
    n = sc.n
    stream = torch.cuda.current_stream()
    log(f"[bench] build n={n}")

    build_wall, build_ev = timed(lambda: sc.build(stream), max(3, args.steps), 2)
    build_mtris = n / (build_ev / 1e3) / 1e6
    build_gbs = BUILD_BYTES_PER_TRI * n / (build_ev / 1e3) / 1e9

    cnt = torch.zeros(3, dtype=torch.int64, device=dev)
    sc.frame(counters=cnt)
    torch.cuda.synchronize()
    rays, fetches, tests = (int(x) for x in cnt.cpu().numpy().view(np.uint64))
    ao_bytes = NODE_BYTES * fetches + TRI_BYTES * tests

    log(f"[bench] AO frame: {rays} rays, {fetches} node fetches, {tests} triangle tests")
    ao_wall, ao_ev = timed(lambda: sc.frame(stream=stream), args.steps, args.warmup)
    mrays = rays / (ao_wall / 1e3) / 1e6
    ao_gbs = ao_bytes / (ao_ev / 1e3) / 1e9

    # The traversal's roof: random 64-B node gathers from an L2-resident
    # table, measured now on this GPU (csrc/calib/probes.hip): per-lane
    # fetches (the traversal's pattern, independent and dependent chains at
    # the AO kernel's residency) and quad-cooperative fetches (the best 64-B
    # gather rate measured); plus the same from a MALL-sized table.
    log("[bench] gather roofline probes")
    gather = {}
    workloads.run_gather(5, gather, probes=[p for p in workloads.GATHER_PROBES if not p[0].startswith("hbm")])
    per_lane_peak = max(gather["l2_independent"]["gbs"], gather["l2_dependent_ao_occupancy"]["gbs"])
    l2_peak = max(per_lane_peak, gather["l2_quad_cooperative"]["gbs"])

    # standalone sort of 2^k uniform u64 keys, 8 passes of 8-bit digits
    N = 1 << args.sort_log2
    log(f"[bench] sort N={N}")
    keys0 = workloads.random_u64(N)
    keys = torch.empty_like(keys0)
    tmp = torch.empty_like(keys0)

    sort_ms = []
    for i in range(args.warmup + args.steps):
        keys.copy_(keys0)
        s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s.record()
        sc.sorter.sort(keys, None, tmp, None, N, 0, 64, stream)
        e.record()
        torch.cuda.synchronize()
        if i >= args.warmup:
            sort_ms.append(s.elapsed_time(e))
    sort_avg = sum(sort_ms) / len(sort_ms)
    # The fraction is taken over the bytes this sort actually moves on its
    # large-input route: one read for the three top digits' histograms, three
    # digit passes (read + write) and the segment finish (read + write) =
    # 9 * 8 * N.  An 8-pass LSD sort would move (2P+1)*8*N with P = 8; at 60 %
    # of peak that is 0.951 ms at N = 2^25, reported as a time bound only.
    P = 8
    run_bytes = (1 + 2 * 3 + 2) * 8 * N
    run_gbs = run_bytes / (sort_avg / 1e3) / 1e9
    del keys0, keys, tmp

    batched = {}
    if not args.no_batched:
        log("[bench] batched build + trace (4096 meshes)")
        workloads.run_batched(max(args.steps, 3), batched)
        torch.cuda.empty_cache()
    walks = {}
    if args.level == 9:
        log("[bench] escape-link / indexed traversal, indexed and 63-bit builds")
        workloads.run_walks(max(args.steps, 3), walks, sc)
    ptres = {}
    if not args.no_pt:
        log("[bench] path tracing: 64 spheres, 21M triangles, 1080p 8 spp")
        workloads.run_pt(3, ptres)
        torch.cuda.empty_cache()
    diffres = {}
    if not args.no_diff:
        log("[bench] differentiable render: 200 Adam steps")
        workloads.run_diff(200, diffres)

    # HBM bytes per k_render_ao_pool launch: FETCH_SIZE x2 + WRITE_SIZE x1 (copy-calibrated) from the
    # rocprofv3 --pmc runs of the same workload, summarised in profiles/
    ao_traffic, ao_traffic_src, ao_l2_hit = None, None, None
    prof = os.path.join(ROOT, "profiles", "k_render_ao_pool_l9_1080p.txt")
    if args.level == 9 and os.path.exists(prof):
        for line in open(prof):
            if line.startswith("traffic_bytes_per_launch:"):
                ao_traffic = int(float(line.split(":", 1)[1]))
                ao_traffic_src = "profiles/k_render_ao_pool_l9_1080p.txt"
            elif line.startswith("L2 hit rate:"):
                ao_l2_hit = float(line.split(":", 1)[1].split()[0])

    out = {
        "metric": METRIC,
        "value": round(mrays, 2),
        "unit": "MRays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ao_wall, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"ao_icosphere_l{args.level}_{W}x{H}_{SPP}spp",
            "triangles": n,
            "width": W,
            "height": H,
            "ao_spp": SPP,
            "ao_radius": RADIUS,
            "sampler": "owen-scrambled sobol, seed 0",
            "rays_per_step": rays,
        },
        "roofline": {
            "kernel": "k_render_ao_pool",
            "bound": "l2_gather",
            "achieved": round(ao_gbs, 1),
            "peak": round(l2_peak, 1),
            "unit": "GB/s",
            "frac": round(ao_gbs / l2_peak, 4),
            "traffic": ao_traffic,
            "traffic_source": ao_traffic_src,
            "bytes_per_launch": ao_bytes,
            "avg_launch_ms": round(ao_ev, 4),
            "peak_source": "random 64-B node gathers from a 2 MiB (L2-resident) table, best of per-lane and "
                           "quad-cooperative fetch, measured in this run (gather.*)",
            "per_lane_fetch_peak": round(per_lane_peak, 1),
            "per_lane_fetch_frac": round(ao_gbs / per_lane_peak, 4),
            "hbm_gbs": round(ao_traffic / (ao_ev / 1e3) / 1e9, 1) if ao_traffic else None,
            "hbm_frac": round(ao_traffic / (ao_ev / 1e3) / 1e9 / HBM_PEAK_GBS, 4) if ao_traffic else None,
            "gather": gather,
            "l2_hit_rate": ao_l2_hit,
            "note": "achieved = algorithmic bytes (64 B x node fetches + 36 B x triangle tests, counts equal to "
                    "the oracle's walks) / kernel time.  Most of the kernel's loads hit L2 (l2_hit_rate, from the "
                    "same profile as traffic), so the roof is the L2 gather rate of 64-B nodes; HBM moves a few "
                    "percent of its peak (hbm_frac)",
        },
        "build": {
            "triangles": n,
            "ms": round(build_ev, 4),
            "mtris_per_s": round(build_mtris, 1),
            "algorithmic_bytes": BUILD_BYTES_PER_TRI * n,
            "achieved_gbs": round(build_gbs, 1),
            "frac": round(build_gbs / HBM_PEAK_GBS, 4),
        },
        "sort": {
            "keys": N,
            "key_bits": 64,
            "ms": round(sort_avg, 4),
            "gkeys_per_s": round(N / (sort_avg / 1e3) / 1e9, 2),
            "route": "3 top-digit onesweep passes (the first reads tile prefixes scanned from the histogram read, "
                     "no look-back) + segment finish (full 8-pass fallback only for long runs of "
                     "equal 24-bit prefix)",
            "algorithmic_bytes": run_bytes,
            "achieved_gbs": round(run_gbs, 1),
            "frac": round(run_gbs / HBM_PEAK_GBS, 4),
            "bytes_model": "bytes the route moves: 2 x 8 B per key for each of the 3 passes and the finish, "
                           "plus the 8-B histogram read",
            "lsd8_equivalent_ms": round((2 * P + 1) * 8 * N / (0.6 * HBM_PEAK_GBS * 1e9) * 1e3, 4),
            "lsd8_equivalent_note": "time an 8-pass LSD sort would need at 60 % of 8 TB/s; the route sorts in "
                                    "fewer passes, so this is a time bound, not a bandwidth fraction",
            "tile": "x".join(str(v) for v in autotune.SORT_CONFIGS[best_cfg]),
        },
        "batched": batched or None,
        "walks": walks or None,
        "pt": ptres or None,
        "diff": diffres or None,
        "tuning": autotune.get_tuning(),
    }

    if not args.no_cpu:
        from oracle import pyoracle as orc

        threads = orc.max_threads()
        log(f"[bench] CPU oracle baseline on {threads} threads")
        # the build in the reference CPU builder's structure (BVHCPUBuilder::build
        # with BUILD_OPTION_CPU_PARALLEL, minimum_lbvh.h:512-591): serial scene
        # box and keys, std::sort on the key alone, leaf-parallel climb
        t0 = orc.now()
        nodes, root, _, _ = orc.build(sc.tris_np, opt=8)
        cpu_build_s = orc.now() - t0
        t0 = orc.now()
        nodes, root, _, _ = orc.build(sc.tris_np)
        cpu_radix_s = orc.now() - t0
        og = orc.raygen_lookat(sc.eye, sc.ctr, sc.up, sc.fovy, W, H)
        t0 = orc.now()
        _, _, ccnt = orc.render_ao(og, W, H, root, nodes, sc.tris_np, 1, SPP, RADIUS, SEED)
        cpu_ao_s = orc.now() - t0
        out["cpu_baseline"] = {
            "value": round(int(ccnt[0]) / cpu_ao_s / 1e6, 3),
            "unit": "MRays/s",
            "cores": threads,
            "kind": "port",
            "sample": f"the whole AO frame ({int(ccnt[0])} rays, {cpu_ao_s:.2f} s on {threads} threads), the "
                      f"reference's stack-free traversal (oracle restatement), OpenMP over rows; plus one build "
                      f"of the {n}-tri mesh in the reference CPU builder's structure",
            "build_mtris_per_s": round(n / cpu_build_s / 1e6, 2),
            "build_s": round(cpu_build_s, 3),
            "build_structure": "serial scene AABB and keys, std::sort on the key (minimum_lbvh.h:554), "
                               "leaf-parallel climb (:565)",
            "build_radix_mtris_per_s": round(n / cpu_radix_s / 1e6, 2),
            "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
        }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
