"""Graph capture and scratch allocation: a build that needs more scratch than
its context holds cannot be captured (a graph cannot record hipMalloc).  The
C ABI then returns MLBVH_E_INVALID_ARG with a message that says what to do,
and a build at a size the context has already served captures and replays
bit for bit."""
import ctypes

import numpy as np
import pytest
import torch

from minimumlbvh_amd import _capi, scenes

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = ctypes.CDLL("libamdhip64.so")
        P = ctypes.c_void_p
        _hip.hipStreamCreate.argtypes = [ctypes.POINTER(P)]
        _hip.hipStreamDestroy.argtypes = [P]
        _hip.hipStreamBeginCapture.argtypes = [P, ctypes.c_int]
        _hip.hipStreamEndCapture.argtypes = [P, ctypes.POINTER(P)]
        _hip.hipGraphDestroy.argtypes = [P]
        _hip.hipGraphInstantiate.argtypes = [ctypes.POINTER(P), P, P, P, ctypes.c_size_t]
        _hip.hipGraphLaunch.argtypes = [P, P]
        _hip.hipGraphExecDestroy.argtypes = [P]
        _hip.hipStreamSynchronize.argtypes = [P]
        _hip.hipGetLastError.argtypes = []
        _hip.hipStreamIsCapturing.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    return _hip


GLOBAL, THREAD_LOCAL = 0, 1  # hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def test_capture_needs_warm_scratch():
    assert torch.cuda.is_available()
    lib, H = _capi.load(), hip()
    n = 200_000
    tris = torch.from_numpy(scenes.random_triangles(n, 21)).cuda()
    nodes = torch.empty((n - 1, 16), dtype=torch.int32, device="cuda")
    root = torch.empty(1, dtype=torch.int32, device="cuda")
    ctx = ctypes.c_void_p()
    assert lib.mlbvh_ctx_create(0, ctypes.byref(ctx)) == _capi.MLBVH_OK
    stream = ctypes.c_void_p()
    assert H.hipStreamCreate(ctypes.byref(stream)) == 0
    torch.cuda.synchronize()
    try:
        # a fresh context has no build scratch: the captured call must refuse
        assert H.hipStreamBeginCapture(stream, GLOBAL) == 0
        rc = lib.mlbvh_build(ctx, _ptr(tris), n, 0, _ptr(nodes), _ptr(root), stream)
        graph = ctypes.c_void_p()
        end = H.hipStreamEndCapture(stream, ctypes.byref(graph))
        if graph.value:
            H.hipGraphDestroy(graph)
        print(f"captured build on a fresh context: rc {rc}, end capture {end}, {lib.mlbvh_last_error()!r}", flush=True)
        assert rc == _capi.MLBVH_E_INVALID_ARG
        assert b"stream is being captured" in lib.mlbvh_last_error()
        # the invalidated capture is the application's to clean up: clear the
        # runtime's last error and continue on another stream (on this ROCm the
        # stream keeps capture status 2, invalidated, after hipStreamEndCapture
        # and later launches on it fail with error 901)
        status = ctypes.c_int(-1)
        print(f"last error {H.hipGetLastError()}, is-capturing {H.hipStreamIsCapturing(stream, ctypes.byref(status))} "
              f"status {status.value}", flush=True)
        H.hipStreamDestroy(stream)
        assert H.hipStreamCreate(ctypes.byref(stream)) == 0
        # warm the scratch outside capture, then capture the same size and replay
        rc = lib.mlbvh_build(ctx, _ptr(tris), n, 0, _ptr(nodes), _ptr(root), stream)
        assert rc == _capi.MLBVH_OK, lib.mlbvh_last_error()
        assert H.hipStreamSynchronize(stream) == 0
        want_nodes, want_root = nodes.clone(), root.clone()
        assert H.hipStreamBeginCapture(stream, THREAD_LOCAL) == 0
        rc = lib.mlbvh_build(ctx, _ptr(tris), n, 0, _ptr(nodes), _ptr(root), stream)
        graph = ctypes.c_void_p()
        assert H.hipStreamEndCapture(stream, ctypes.byref(graph)) == 0
        assert rc == _capi.MLBVH_OK, lib.mlbvh_last_error()
        exe = ctypes.c_void_p()
        assert H.hipGraphInstantiate(ctypes.byref(exe), graph, None, None, 0) == 0
        for _ in range(3):
            nodes.fill_(-1)
            root.fill_(-1)
            torch.cuda.synchronize()
            assert H.hipGraphLaunch(exe, stream) == 0
            assert H.hipStreamSynchronize(stream) == 0
            assert torch.equal(nodes, want_nodes) and torch.equal(root, want_root)
        H.hipGraphExecDestroy(exe)
        H.hipGraphDestroy(graph)
    finally:
        H.hipStreamSynchronize(stream)
        H.hipStreamDestroy(stream)
        lib.mlbvh_ctx_destroy(ctx)
    assert np.asarray(want_root.cpu())[0] >= 0
