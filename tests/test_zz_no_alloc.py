"""mlbvh_build_ex_noalloc through the C ABI: the same nodes, root, sorted
records and deltas as mlbvh_build_ex, bit for bit, on two calls in a row (the
second with a smaller input, so the scratch holds the first call's larger
contents), for every build option; no device memory is allocated by the
call; sizes above the limit are rejected."""
import ctypes

import numpy as np
import pytest
import torch

from minimumlbvh_amd import _capi, scenes

pytestmark = pytest.mark.gpu

MAX_TRIS = 1 << 20


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _build(lib, ctx, fn, tris, opt):
    n = tris.shape[0]
    dev = tris.device
    nodes = torch.full((n - 1, 16), -1, dtype=torch.int32, device=dev)
    root = torch.full((1,), -1, dtype=torch.int32, device=dev)
    rec = torch.zeros(n, dtype=torch.int64, device=dev)
    deltas = torch.zeros(n - 1, dtype=torch.uint8, device=dev)
    rc = getattr(lib, fn)(ctx, _ptr(tris), n, opt, _ptr(nodes), _ptr(root), _ptr(rec), _ptr(deltas), None)
    assert rc == _capi.MLBVH_OK, lib.mlbvh_last_error()
    torch.cuda.synchronize()
    return nodes, root, rec, deltas


@pytest.fixture(scope="module")
def capi():
    assert torch.cuda.is_available(), "gpu tests need a ROCm GPU"
    lib = _capi.load()
    ctx = ctypes.c_void_p()
    assert lib.mlbvh_ctx_create(0, ctypes.byref(ctx)) == _capi.MLBVH_OK, lib.mlbvh_last_error()
    yield lib, ctx
    lib.mlbvh_ctx_destroy(ctx)


@pytest.mark.parametrize("opt", [0, 1, 4])
def test_noalloc_build_equals_build_ex_twice(capi, opt):
    lib, ctx = capi
    for n, seed in ((MAX_TRIS, 5 + opt), (300_001, 6 + opt), (1000, 7 + opt)):
        print(f"noalloc build opt={opt} n={n}", flush=True)
        tris = torch.from_numpy(scenes.random_triangles(n, seed)).cuda()
        want = _build(lib, ctx, "mlbvh_build_ex", tris, opt)
        got = _build(lib, ctx, "mlbvh_build_ex_noalloc", tris, opt)
        for w, g in zip(want, got):
            assert torch.equal(w, g)


def test_noalloc_build_allocates_nothing(capi):
    lib, ctx = capi
    tris = torch.from_numpy(scenes.random_triangles(500_000, 11)).cuda()
    n = tris.shape[0]
    nodes = torch.empty((n - 1, 16), dtype=torch.int32, device="cuda")
    root = torch.empty(1, dtype=torch.int32, device="cuda")
    fresh = ctypes.c_void_p()
    assert lib.mlbvh_ctx_create(0, ctypes.byref(fresh)) == _capi.MLBVH_OK
    try:
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info(0)
        for _ in range(3):
            rc = lib.mlbvh_build_ex_noalloc(fresh, _ptr(tris), n, 0, _ptr(nodes), _ptr(root), None, None, None)
            assert rc == _capi.MLBVH_OK, lib.mlbvh_last_error()
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info(0)
        # a fresh context's plain build would allocate about 40 MB of scratch here
        assert free1 == free0
        want_nodes, want_root, _, _ = _build(lib, fresh, "mlbvh_build_ex", tris, 0)
        assert torch.equal(nodes, want_nodes) and torch.equal(root, want_root)
    finally:
        lib.mlbvh_ctx_destroy(fresh)


def test_noalloc_build_rejects_sizes_above_the_limit(capi):
    lib, ctx = capi
    rc = lib.mlbvh_build_ex_noalloc(ctx, ctypes.c_void_p(16), MAX_TRIS + 1, 0, ctypes.c_void_p(16),
                                    ctypes.c_void_p(16), None, None, None)
    assert rc == _capi.MLBVH_E_INVALID_ARG
    assert b"maximum" in lib.mlbvh_last_error()
    assert lib.mlbvh_build_ex_noalloc(ctx, None, 100, 0, None, None, None, None, None) == _capi.MLBVH_E_INVALID_ARG
    assert lib.mlbvh_build_ex_noalloc(None, None, 100, 0, None, None, None, None, None) == _capi.MLBVH_E_INVALID_ARG
