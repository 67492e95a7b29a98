"""Batched build + persistent trace (mlbvh_build_batched / mlbvh_trace_batched):
every mesh's BVH is bit-identical to building it alone (oracle and the plain
GPU path), and batched hits are bit-identical to the oracle's per-mesh
traces."""
import numpy as np
import pytest
import torch

from minimumlbvh_amd import scenes

pytestmark = pytest.mark.gpu


def _mesh_nodes(nodes, toff, m):
    return nodes[toff[m] - m: toff[m + 1] - m - 1]


def _run_batched(gpu, tris, toff, rays, opt=0, query=0):
    bb = gpu.BatchedBVH(toff)
    t = torch.from_numpy(tris).cuda()
    bb.build(t, opt)
    r = torch.from_numpy(rays).cuda()
    hits = bb.trace(t, r, query)
    torch.cuda.synchronize()
    return bb, t, r, bb.nodes.cpu().numpy().view(np.uint32), bb.roots.cpu().numpy().view(np.uint32), \
        hits.cpu().numpy().view(np.uint32).reshape(-1, 8)


@pytest.mark.parametrize("opt", [0, 1])
def test_batched_build_and_trace_match_oracle(gpu, orc, opt):
    counts = np.array([2, 3, 257, 4096, 65, 1000, 2, 7777, 300, 12000])
    rcounts = np.array([100, 50, 400, 1000, 10, 800, 5, 1200, 300, 2000])
    tris, toff, rays, roff = scenes.batch_scene(counts, rcounts, seed=3)
    _, _, _, nodes, roots, hits = _run_batched(gpu, tris, toff, rays, opt)
    for m in range(len(counts)):
        mt = tris[toff[m]:toff[m + 1]]
        on, oroot, _, _ = orc.build(mt, opt=opt)
        assert roots[m] == oroot, m
        assert np.array_equal(_mesh_nodes(nodes, toff, m), on), m
        mr = rays[roff[m]:roff[m + 1]].copy()
        mr[:, 7] = 0
        oh, _ = orc.trace(on, mt, oroot, mr)
        assert np.array_equal(hits[roff[m]:roff[m + 1]], oh.view(np.uint32).reshape(-1, 8)), m


def test_batched_equals_plain_per_mesh_path(gpu):
    tc, rc = scenes.batch_sizes(300, seed=5, tri_range=(64, 20000), ray_range=(64, 4000))
    tris, toff, rays, roff = scenes.batch_scene(tc, rc, seed=5)
    bb, t, r, nodes, roots, hits = _run_batched(gpu, tris, toff, rays)
    bb.nodes.zero_()
    bb.roots.zero_()
    phits = bb.build_trace_per_mesh(t, r, roff)
    torch.cuda.synchronize()
    assert np.array_equal(bb.roots.cpu().numpy().view(np.uint32), roots)
    assert np.array_equal(bb.nodes.cpu().numpy().view(np.uint32), nodes)
    assert np.array_equal(phits.cpu().numpy().view(np.uint32).reshape(-1, 8), hits)
    assert (hits[:, 7] != 0xFFFFFFFF).mean() > 0.05


def test_batched_any_hit_and_tuning(gpu, orc):
    counts = np.array([500, 20, 9000])
    rcounts = np.array([3000, 40, 5000])
    tris, toff, rays, roff = scenes.batch_scene(counts, rcounts, seed=8)
    rays[:, 3] = 1.5
    ref = None
    for bpc, chunk in [(8, 64), (2, 256), (16, 128)]:
        gpu.set_tuning(2, bpc, chunk)
        _, _, _, _, _, hits = _run_batched(gpu, tris, toff, rays, query=1)
        if ref is None:
            ref = hits
        assert np.array_equal(hits, ref)
    gpu.set_tuning(2, 8, 64)
    for m in range(3):
        mt = tris[toff[m]:toff[m + 1]]
        on, oroot, _, _ = orc.build(mt)
        mr = rays[roff[m]:roff[m + 1]].copy()
        mr[:, 7] = 0
        oh, _ = orc.trace(on, mt, oroot, mr, query=1)
        assert np.array_equal(ref[roff[m]:roff[m + 1]], oh.view(np.uint32).reshape(-1, 8))


def test_batched_ray_naming_no_mesh_misses(gpu):
    """A ray whose mesh id is not a mesh of the batch gets a miss record
    (t = its tmax, uv and ng zero, triangle 0xFFFFFFFF); the other rays'
    hits are unchanged."""
    counts = np.array([300, 5000, 64])
    rcounts = np.array([700, 900, 300])
    tris, toff, rays, roff = scenes.batch_scene(counts, rcounts, seed=13)
    _, _, _, _, _, ref = _run_batched(gpu, tris, toff, rays)
    bad = rays.copy()
    idx = np.arange(0, bad.shape[0], 7)
    mesh = bad[:, 7].view(np.uint32)
    mesh[idx] = np.where(idx % 2, 3, 0xFFFFFFFF).astype(np.uint32)
    bad[idx[::3], 3] = 0.75
    _, _, _, _, _, hits = _run_batched(gpu, tris, toff, bad)
    keep = np.ones(bad.shape[0], bool)
    keep[idx] = False
    assert np.array_equal(hits[keep], ref[keep])
    miss = np.zeros((idx.size, 8), np.uint32)
    miss[:, 0] = bad[idx, 3].view(np.uint32)
    miss[:, 7] = 0xFFFFFFFF
    assert np.array_equal(hits[idx], miss)


def test_batched_rejects_bad_offsets(gpu):
    with pytest.raises(ValueError):
        gpu.BatchedBVH([0, 5, 6])  # a one-triangle mesh


def test_batched_build_trace_graph_replays(gpu):
    """build -> trace captured in one graph: every replay, with eager work of
    the same library (a standalone sort, a plain build) and clobbered outputs
    in between, reproduces the eager results bit for bit."""
    tc, rc = scenes.batch_sizes(64, seed=11, tri_range=(64, 30000), ray_range=(64, 3000))
    tris, toff, rays, roff = scenes.batch_scene(tc, rc, seed=11)
    bb, t, r, nodes, roots, hits = _run_batched(gpu, tris, toff, rays)
    out = torch.empty((rays.shape[0], 8), dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        bb.build(t, 0, s)
        bb.trace(t, r, hits=out, stream=s)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        bb.build(t)
        bb.trace(t, r, hits=out)
    keys = torch.randint(-(2**63), 2**63 - 1, (100_000,), dtype=torch.int64, device="cuda")
    keys_np = keys.cpu().numpy().view(np.uint64)
    ktmp = torch.empty_like(keys)
    plain = gpu.BVHGPUBuilder()
    sorter = gpu.OnesweepSort()
    for rep in range(4):
        print(f"graph replay {rep}", flush=True)
        out.fill_(-1)
        bb.nodes.fill_(-1)
        bb.roots.fill_(-1)
        k = keys.clone()
        sorter.sort(k, None, ktmp, None, k.numel(), 0, 64)
        plain.build(t[: 5000 + 1000 * rep].contiguous(), 5000 + 1000 * rep, 0, sorter)
        torch.cuda.synchronize()
        assert np.array_equal(k.cpu().numpy().view(np.uint64), np.sort(keys_np))
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().view(np.uint32).reshape(-1, 8), hits), rep
        assert np.array_equal(bb.nodes.cpu().numpy().view(np.uint32), nodes), rep
        assert np.array_equal(bb.roots.cpu().numpy().view(np.uint32), roots), rep


def test_batched_full_configuration(gpu, orc):
    """The batching configuration at full size (4096 meshes, about 1.3e8
    triangles and 4.8e7 rays, generated on the device): one segmented build +
    persistent trace gives the same nodes, roots and hits, bit for bit, as the
    plain per-mesh build + trace loop, and, for the six largest meshes (up to
    262,144 triangles) and six others drawn at random, as the oracle's build
    and trace of that mesh alone."""
    from minimumlbvh_amd import workloads

    sc = workloads.BatchedScene()
    print(f"batched full size: {sc.n_meshes} meshes, {sc.n_tris} triangles, {sc.n_rays} rays", flush=True)
    sc.build_trace()
    torch.cuda.synchronize()
    nodes, roots, hits = sc.bvh.nodes.clone(), sc.bvh.roots.clone(), sc.hits.clone()
    sc.bvh.nodes.fill_(-1)
    sc.bvh.roots.fill_(-1)
    loop_hits = sc.per_mesh_loop()
    torch.cuda.synchronize()
    assert torch.equal(sc.bvh.roots, roots)
    assert torch.equal(sc.bvh.nodes, nodes)
    assert torch.equal(loop_hits, hits)
    assert (hits[:, 7] != -1).float().mean().item() > 0.05
    # against the oracle, mesh by mesh, at the sizes only the full batch has
    sizes = np.diff(sc.toff.astype(np.int64))
    pick = list(np.argsort(sizes)[-6:]) + list(np.random.default_rng(4).choice(sc.n_meshes, 6, replace=False))
    roots_np = roots.cpu().numpy().view(np.uint32)
    for m in pick:
        t0, t1 = int(sc.toff[m]), int(sc.toff[m + 1])
        r0, r1 = int(sc.roff[m]), int(sc.roff[m + 1])
        mt = sc.tris[t0:t1].cpu().numpy()
        on, oroot, _, _ = orc.build(mt)
        print(f"batched full size vs oracle: mesh {m}, {t1 - t0} triangles, {r1 - r0} rays", flush=True)
        assert roots_np[m] == oroot, m
        assert np.array_equal(nodes[t0 - m: t1 - m - 1].cpu().numpy().view(np.uint32), on), m
        mr = sc.rays[r0:r1].cpu().numpy().copy()
        mr[:, 7] = 0
        oh, _ = orc.trace(on, mt, oroot, mr)
        assert np.array_equal(hits[r0:r1].cpu().numpy().view(np.uint32), oh.view(np.uint32).reshape(-1, 8)), m
    assert sizes[pick[5]] > 250_000
