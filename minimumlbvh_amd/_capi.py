"""ctypes binding of libmlbvh.so (the C ABI declared in include/mlbvh.h).

The library is required: if it is missing or cannot be loaded this module
raises, there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, c_char_p, c_float, c_int, c_uint32, c_ulonglong, c_void_p

# torch first: libmlbvh must bind to the HIP runtime that torch has loaded.
# Loaded the other way round, a process can end up with two HIP runtimes and
# the second one sees no device.
import torch  # noqa: F401

from ._build import LIB as LIB_PATH

MLBVH_OK = 0
MLBVH_E_INVALID_ARG = -1
MLBVH_E_HIP = -2
MLBVH_E_NO_DEVICE = -3

BUILD_OPTION_USE_NORMAL = 1
BUILD_OPTION_CPU_PARALLEL = 2
BUILD_OPTION_MORTON64 = 4
RAY_QUERY_CLOSEST = 0
RAY_QUERY_ANY = 1

# Every exported entry point and its ctypes signature (argtypes, restype).
_P = c_void_p
SIGNATURES = {
    "mlbvh_abi_version": ([], c_int),
    "mlbvh_last_error": ([], c_char_p),
    "mlbvh_ctx_create": ([c_int, POINTER(c_void_p)], c_int),
    "mlbvh_ctx_destroy": ([_P], c_int),
    "mlbvh_sort_u32": ([_P, _P, _P, _P, _P, c_uint32, c_int, c_int, _P], c_int),
    "mlbvh_sort_u64": ([_P, _P, _P, _P, _P, c_uint32, c_int, c_int, _P], c_int),
    "mlbvh_build": ([_P, _P, c_int, c_uint32, _P, _P, _P], c_int),
    "mlbvh_build_ex": ([_P, _P, c_int, c_uint32, _P, _P, _P, _P, _P], c_int),
    "mlbvh_build_ex_noalloc": ([_P, _P, c_int, c_uint32, _P, _P, _P, _P, _P], c_int),
    "mlbvh_refit": ([_P, _P, c_int, _P, _P, _P, _P, _P], c_int),
    "mlbvh_build_indexed": ([_P, _P, c_int, _P, c_int, c_uint32, _P, _P, _P, _P, _P], c_int),
    "mlbvh_refit_indexed": ([_P, _P, c_int, _P, c_int, _P, _P, _P, _P, _P], c_int),
    "mlbvh_trace": ([_P, _P, _P, _P, _P, _P, c_int, c_int, _P, _P], c_int),
    "mlbvh_trace_escape_link": ([_P, _P, _P, _P, _P, _P, c_int, c_int, _P, _P], c_int),
    "mlbvh_trace_stackfree": ([_P, _P, _P, _P, _P, _P, c_int, c_int, _P, _P], c_int),
    "mlbvh_trace_indexed": ([_P, _P, _P, c_int, _P, _P, _P, _P, c_int, c_int, _P, _P], c_int),
    "mlbvh_build_batched": ([_P, _P, _P, c_int, c_uint32, c_uint32, c_uint32, _P, _P, _P], c_int),
    "mlbvh_trace_batched": ([_P, _P, _P, _P, _P, c_int, _P, _P, c_int, c_int, _P, _P], c_int),
    "mlbvh_build_trace_per_mesh": ([_P, _P, _P, c_int, c_uint32, _P, _P, _P, _P, _P, c_int, _P], c_int),
    "mlbvh_set_tuning": ([_P, c_int, c_int, c_int], c_int),
    "mlbvh_get_tuning": ([_P, _P, _P, _P], c_int),
    "mlbvh_gather_triangles": ([_P, _P, _P, c_int, _P, _P], c_int),
    "mlbvh_render_primary": ([_P, _P, _P, c_int, c_int, _P, POINTER(c_float), _P, _P, _P, _P], c_int),
    "mlbvh_primary_adjoint": ([_P, _P, _P, c_int, c_int, _P, POINTER(c_float), _P, _P, _P, _P], c_int),
    "mlbvh_raygen_lookat": ([_P, POINTER(c_float), POINTER(c_float), POINTER(c_float), c_float, c_int, c_int], c_int),
    "mlbvh_render_normal": ([_P, _P, c_int, c_int, _P, _P, _P, _P, _P], c_int),
    "mlbvh_render_ao": ([_P, _P, _P, c_int, c_int, _P, _P, _P, _P, c_int, c_int, c_float, c_uint32, _P, _P], c_int),
    "mlbvh_render_pt": ([_P, _P, c_int, c_int, _P, _P, _P, _P, _P, c_int, c_int, c_int, _P], c_int),
    "mlbvh_pack": ([_P, _P, _P, c_int, _P], c_int),
}


class MlbvhError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"mlbvh status {status}: {message}")
        self.status = status


class RayGen(ctypes.Structure):
    """RayGenerator of camera.h: origin, right, up (36 bytes)."""

    _fields_ = [("origin", c_float * 3), ("right", c_float * 3), ("up", c_float * 3)]


_lib = None


def load(path: str | None = None) -> ctypes.CDLL:
    """Load libmlbvh.so; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"libmlbvh.so not found at {p}; build it with `python -m minimumlbvh_amd._build` "
            "(the HIP library is required, there is no CPU fallback)"
        )
    lib = ctypes.CDLL(p)
    for name, (argtypes, restype) in SIGNATURES.items():
        # an entry point missing from an older build fails when it is called
        # (AttributeError); tests/test_capi.py requires all of them
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.argtypes = argtypes
        fn.restype = restype
    if lib.mlbvh_abi_version() != 1:
        raise ImportError(f"libmlbvh ABI {lib.mlbvh_abi_version()} != 1")
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != MLBVH_OK:
        msg = load().mlbvh_last_error()
        raise MlbvhError(status, msg.decode() if msg else "")
