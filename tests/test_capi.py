"""The C ABI library without a GPU: it loads, exports every symbol that
include/mlbvh.h declares, rejects bad arguments before touching a device, and
its host-side camera (RayGenerator::lookat, camera.h:10-22) is bit-identical
to the oracle."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

from minimumlbvh_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            src = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"MLBVH_API\s+[\w\s\*]+?\b(mlbvh_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    lib = _capi.load()
    names = declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert names == set(_capi.SIGNATURES), "ctypes signatures cover the header"
    assert lib.mlbvh_abi_version() == 1


def test_invalid_arguments_are_rejected_without_a_device():
    lib = _capi.load()
    assert lib.mlbvh_sort_u32(None, None, None, None, None, 10, 0, 32, None) == _capi.MLBVH_E_INVALID_ARG
    assert lib.mlbvh_build(None, None, 10, 0, None, None, None) == _capi.MLBVH_E_INVALID_ARG
    assert lib.mlbvh_trace(None, None, None, None, None, None, 1, 0, None, None) == _capi.MLBVH_E_INVALID_ARG
    assert b"null" in lib.mlbvh_last_error()
    E = _capi.MLBVH_E_INVALID_ARG
    assert lib.mlbvh_trace_stackfree(None, None, None, None, None, None, 1, 0, None, None) == E
    assert lib.mlbvh_trace_escape_link(None, None, None, None, None, None, 1, 0, None, None) == E
    # indexed entry points check their vertex / index buffers before any device work
    assert lib.mlbvh_build_indexed(None, None, 3, None, 10, 0, None, None, None, None, None) == E
    assert b"build_indexed" in lib.mlbvh_last_error()
    assert lib.mlbvh_refit_indexed(None, None, 0, None, 10, None, None, None, None, None) == E
    assert lib.mlbvh_trace_indexed(None, None, None, 0, None, None, None, None, 1, 0, None, None) == E
    assert b"trace_indexed" in lib.mlbvh_last_error()
    out = ctypes.c_void_p()
    assert lib.mlbvh_ctx_create(0, None) == _capi.MLBVH_E_INVALID_ARG
    rc = lib.mlbvh_ctx_create(-1, ctypes.byref(out))
    assert rc == _capi.MLBVH_E_NO_DEVICE and not out.value


def test_lookat_bit_identical_to_oracle(orc):
    from minimumlbvh_amd.lbvh import RayGenerator

    for eye, ctr, up, fov, w, h in [((0.5, 0.5, -1.5), (0.5, 0.5, -0.5), (0, 1, 0), 60.0, 256, 256),
                                    ((0.0, 0.3, -2.6), (0, 0, 0), (0, 1, 0), 60.0, 1920, 1080),
                                    ((1.0, 2.0, 3.0), (-1.0, 0.5, 0.25), (0, 1, 0), 37.0, 640, 480)]:
        g = RayGenerator().lookat(eye, ctr, up, math.radians(fov), w, h)
        o = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
        a = np.array(list(g.g.origin) + list(g.g.right) + list(g.g.up), np.float32)
        b = np.array(list(o.origin) + list(o.right) + list(o.up), np.float32)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_lookat_rejects_bad_size():
    lib = _capi.load()
    g = _capi.RayGen()
    f3 = ctypes.c_float * 3
    rc = lib.mlbvh_raygen_lookat(ctypes.byref(g), f3(0, 0, 0), f3(0, 0, 1), f3(0, 1, 0), 1.0, 0, 10)
    assert rc == _capi.MLBVH_E_INVALID_ARG


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _capi.load(str(tmp_path / "nope.so"))
