"""Indexed-mesh input (mlbvh_build_indexed / mlbvh_refit_indexed /
mlbvh_trace_indexed): u32 index triples into fp32 vertices, as the icosphere
workloads are authored.  Bar: nodes, root, sorted records and deltas
bit-identical to the build of the de-indexed triangle soup (and so to the
oracle); hits and visit counters bit-identical to the soup trace."""
import numpy as np
import pytest
import torch

from minimumlbvh_amd import scenes
from tests.test_gpu_trace import random_rays

pytestmark = pytest.mark.gpu


def _indexed(gpu, v, f, opt=0, keep=False):
    vt = torch.from_numpy(np.ascontiguousarray(v, np.float32)).cuda()
    ft = torch.from_numpy(np.ascontiguousarray(f).astype(np.uint32).view(np.int32)).cuda()
    b = gpu.BVHGPUBuilder()
    b.build_indexed(vt, ft, opt, keep_topology=keep)
    torch.cuda.synchronize()
    return b, vt, ft


def _soup(gpu, tris, opt=0, keep=False):
    t = torch.from_numpy(np.ascontiguousarray(tris, np.float32)).cuda()
    b = gpu.BVHGPUBuilder()
    b.build(t, tris.shape[0], opt, keep_topology=keep)
    torch.cuda.synchronize()
    return b, t


def _arrays(b):
    return (b.m_internals.cpu().numpy().view(np.uint32), int(b.m_rootNode.cpu().numpy().view(np.uint32)[0]))


@pytest.mark.parametrize("level,opt", [(3, 0), (6, 0), (6, 1)])
def test_indexed_build_matches_oracle(gpu, orc, level, opt):
    v, f = scenes.displaced_icosphere(level, seed=level)
    b, _, _ = _indexed(gpu, v, f, opt, keep=True)
    gn, groot = _arrays(b)
    on, oroot, rec, d = orc.build(scenes.deindex(v, f), opt=opt)
    assert groot == oroot and np.array_equal(gn, on)
    assert np.array_equal(b.m_sortedRecords.cpu().numpy().view(np.uint64), rec)
    assert np.array_equal(b.m_deltas.cpu().numpy(), d)


def test_indexed_build_level9_equals_soup_build(gpu):
    """The 5,242,880-triangle icosphere from its index buffer: the same nodes
    as building the 36-byte de-indexed copy."""
    v, f = scenes.displaced_icosphere(9, seed=0)
    bi, _, _ = _indexed(gpu, v, f)
    ni, ri = _arrays(bi)
    del bi
    bs, _ = _soup(gpu, scenes.deindex(v, f))
    ns, rs = _arrays(bs)
    assert ri == rs and np.array_equal(ni, ns)


def test_indexed_refit_matches_soup_refit(gpu, orc):
    v, f = scenes.displaced_icosphere(6, seed=4)
    bi, vt, ft = _indexed(gpu, v, f, keep=True)
    order = (bi.m_sortedRecords.cpu().numpy().view(np.uint64) >> np.uint64(32)).astype(np.uint32)
    deltas = bi.m_deltas.cpu().numpy()
    moved = (v + np.random.default_rng(3).normal(0, 0.002, v.shape)).astype(np.float32)
    vt.copy_(torch.from_numpy(moved).cuda())
    bi.refit_indexed(vt, ft)
    torch.cuda.synchronize()
    gn, groot = _arrays(bi)
    en, eroot = orc.build_from_deltas(deltas, tri_index=order, tris=scenes.deindex(moved, f))
    assert groot == eroot and np.array_equal(gn, en)
    with pytest.raises(ValueError):
        bi.refit_indexed(vt, ft[:-1].contiguous())


@pytest.mark.parametrize("query", [0, 1])
def test_indexed_trace_matches_soup_trace(gpu, orc, query):
    v, f = scenes.displaced_icosphere(6, seed=8)
    tris = scenes.deindex(v, f)
    b, vt, ft = _indexed(gpu, v, f)
    rays = random_rays(60_000, 9, tmax=None if query == 0 else 0.3, lo=-1.5, hi=1.5)
    rays[:, 4:7] = rng_targets(rays, 9) - rays[:, 0:3]
    cnt = torch.zeros(2, dtype=torch.int64, device="cuda")
    gh = gpu.trace_indexed(b, vt, ft, torch.from_numpy(rays).cuda(), query=query, counters=cnt)
    torch.cuda.synchronize()
    nodes, root, _, _ = orc.build(tris)
    oh, _ = orc.trace(nodes, tris, root, rays, query=query)
    _, oc = orc.trace(nodes, tris, root, rays, query=query, method=4)
    got = gh.cpu().numpy().view(np.uint32).reshape(-1, 8)
    assert np.array_equal(got, oh.view(np.uint32).reshape(-1, 8))
    assert np.array_equal(cnt.cpu().numpy().view(np.uint64), oc)
    assert (got[:, 7] != 0xFFFFFFFF).mean() > 0.2


def rng_targets(rays, seed):
    """Points on the unit sphere's neighbourhood, so most rays hit the mesh."""
    d = np.random.default_rng(seed + 1).normal(size=(rays.shape[0], 3))
    return (0.9 * d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def test_indexed_out_of_range_indices_are_clamped(gpu):
    v, f = scenes.displaced_icosphere(4, seed=1)
    bad = f.astype(np.uint32).copy()
    pick = np.random.default_rng(5).choice(bad.size, 50, replace=False)
    bad.flat[pick] = np.uint32(v.shape[0] + 7) + np.arange(50, dtype=np.uint32)
    bad.flat[pick[:5]] = np.uint32(0xFFFFFFFF)
    clamped = np.minimum(bad, np.uint32(v.shape[0] - 1))
    b1, _, _ = _indexed(gpu, v, bad)
    b2, _, _ = _indexed(gpu, v, clamped)
    n1, r1 = _arrays(b1)
    n2, r2 = _arrays(b2)
    assert r1 == r2 and np.array_equal(n1, n2)


def test_indexed_rejects_bad_shapes(gpu):
    v, f = scenes.displaced_icosphere(2, seed=1)
    vt = torch.from_numpy(v).cuda()
    ft = torch.from_numpy(f.astype(np.int32)).cuda()
    b = gpu.BVHGPUBuilder()
    with pytest.raises(ValueError):
        b.build_indexed(vt[:, :2].contiguous(), ft)
    with pytest.raises(ValueError):
        b.build_indexed(vt, ft.to(torch.int64))
    with pytest.raises(ValueError):
        b.build_indexed(vt, ft[:1].contiguous())
