"""Python host API over libmlbvh, mirroring the reference's C++ classes.

Names follow Ushio/MinimumLBVH: ``OnesweepSort`` (tinyhiponesweep.h:572-748),
``BVHGPUBuilder`` (minimum_lbvh.h:726-885) with public ``m_internals`` /
``m_rootNode`` device buffers, ``RayGenerator`` (camera.h), and the render
kernels of main_gpu.cu.  Device memory and streams come from PyTorch-ROCm
tensors; all compute runs in the HIP library.
"""
from __future__ import annotations

import ctypes
import math
from typing import Optional

import torch

from . import _capi
from ._capi import BUILD_OPTION_MORTON64, BUILD_OPTION_USE_NORMAL, RAY_QUERY_ANY, RAY_QUERY_CLOSEST, RayGen, check

NODE_BYTES = 64
TRI_BYTES = 36
HIT_BYTES = 32
INVALID_NODE = 0x7FFFFFFF
LEAF_BIT = 0x80000000

_contexts: dict[int, "Context"] = {}


def _ptr(t: Optional[torch.Tensor]):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def _stream(stream) -> ctypes.c_void_p:
    if stream is None:
        stream = torch.cuda.current_stream()
    if isinstance(stream, torch.cuda.Stream):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


class Context:
    """Per-device library context (sorter + builder scratch)."""

    def __init__(self, device: int = 0):
        self.lib = _capi.load()
        self.device = device
        h = ctypes.c_void_p()
        check(self.lib.mlbvh_ctx_create(device, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            check(self.lib.mlbvh_ctx_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def context(device: int = 0) -> Context:
    if device not in _contexts:
        _contexts[device] = Context(device)
    return _contexts[device]


class OnesweepSort:
    """tinyhiponesweep::OnesweepSort: stable LSD radix sort on the GPU."""

    def __init__(self, device: int = 0):
        self.ctx = context(device)

    def sort(self, keys: torch.Tensor, values: Optional[torch.Tensor], keys_tmp: torch.Tensor,
             values_tmp: Optional[torch.Tensor], n: int, startBit: int, endBit: int, stream=None) -> None:
        lib = self.ctx.lib
        es = keys.element_size()
        fn = {4: lib.mlbvh_sort_u32, 8: lib.mlbvh_sort_u64}.get(es)
        if fn is None:
            raise ValueError("keys must be 32- or 64-bit")
        if keys_tmp.element_size() != es or keys.numel() < n or keys_tmp.numel() < n:
            raise ValueError("keys / keys_tmp size mismatch")
        if values is not None and (values.element_size() != 4 or values.numel() < n or values_tmp is None
                                   or values_tmp.numel() < n):
            raise ValueError("values must be 32-bit with a temp of the same size")
        check(fn(self.ctx.handle, _ptr(keys), _ptr(values), _ptr(keys_tmp), _ptr(values_tmp), n, startBit, endBit,
                 _stream(stream)))


class BVHGPUBuilder:
    """minimum_lbvh::BVHGPUBuilder.  ``kernel``/``includeDir`` are accepted for
    signature compatibility; the kernels are compiled ahead of time."""

    def __init__(self, kernel: Optional[str] = None, includeDir: Optional[str] = None, device: int = 0):
        self.ctx = context(device)
        self.device = device
        self.m_internals: Optional[torch.Tensor] = None  # (n-1, 16) int32 = InternalNode[n-1]
        self.m_rootNode = torch.zeros(1, dtype=torch.int32, device=f"cuda:{device}")
        self.m_nTriangles = 0
        self.m_sortedRecords: Optional[torch.Tensor] = None
        self.m_deltas: Optional[torch.Tensor] = None

    def build(self, triangles: torch.Tensor, nTriangles: int, buildOption: int = 0,
              sorter: Optional[OnesweepSort] = None, stream=None, keep_topology: bool = False) -> None:
        if nTriangles < 2:
            raise ValueError("at least 2 triangles are required")
        if triangles.dtype != torch.float32 or triangles.numel() < 9 * nTriangles:
            raise ValueError("triangles must be float32 with 9 values per triangle")
        dev = triangles.device
        if self.m_internals is None or self.m_internals.shape[0] != nTriangles - 1:
            self.m_internals = torch.empty((nTriangles - 1, 16), dtype=torch.int32, device=dev)
        self.m_nTriangles = nTriangles
        lib = self.ctx.lib
        if keep_topology:
            if self.m_sortedRecords is None or self.m_sortedRecords.numel() != nTriangles:
                self.m_sortedRecords = torch.empty(nTriangles, dtype=torch.int64, device=dev)
                self.m_deltas = torch.empty(nTriangles - 1, dtype=torch.uint8, device=dev)
            check(lib.mlbvh_build_ex(self.ctx.handle, _ptr(triangles), nTriangles, buildOption,
                                     _ptr(self.m_internals), _ptr(self.m_rootNode), _ptr(self.m_sortedRecords),
                                     _ptr(self.m_deltas), _stream(stream)))
        else:
            check(lib.mlbvh_build(self.ctx.handle, _ptr(triangles), nTriangles, buildOption, _ptr(self.m_internals),
                                  _ptr(self.m_rootNode), _stream(stream)))

    def refit(self, triangles: torch.Tensor, stream=None) -> None:
        if self.m_sortedRecords is None:
            raise RuntimeError("refit needs a build(..., keep_topology=True) first")
        check(self.ctx.lib.mlbvh_refit(self.ctx.handle, _ptr(triangles), self.m_nTriangles,
                                       _ptr(self.m_sortedRecords), _ptr(self.m_deltas), _ptr(self.m_internals),
                                       _ptr(self.m_rootNode), _stream(stream)))

    def build_indexed(self, vertices: torch.Tensor, indices: torch.Tensor, buildOption: int = 0, stream=None,
                      keep_topology: bool = False) -> None:
        """Build from an indexed mesh (mlbvh_build_indexed): ``vertices`` (V, 3)
        float32, ``indices`` (T, 3) int32 holding u32 vertex indices.  The nodes
        equal build() on the de-indexed triangles."""
        nV, nT = _check_indexed(vertices, indices)
        if nT < 2:
            raise ValueError("at least 2 triangles are required")
        dev = vertices.device
        if self.m_internals is None or self.m_internals.shape[0] != nT - 1:
            self.m_internals = torch.empty((nT - 1, 16), dtype=torch.int32, device=dev)
        self.m_nTriangles = nT
        if keep_topology:
            if self.m_sortedRecords is None or self.m_sortedRecords.numel() != nT:
                self.m_sortedRecords = torch.empty(nT, dtype=torch.int64, device=dev)
                self.m_deltas = torch.empty(nT - 1, dtype=torch.uint8, device=dev)
        rec = self.m_sortedRecords if keep_topology else None
        dl = self.m_deltas if keep_topology else None
        check(self.ctx.lib.mlbvh_build_indexed(self.ctx.handle, _ptr(vertices), nV, _ptr(indices), nT, buildOption,
                                               _ptr(self.m_internals), _ptr(self.m_rootNode), _ptr(rec), _ptr(dl),
                                               _stream(stream)))

    def refit_indexed(self, vertices: torch.Tensor, indices: torch.Tensor, stream=None) -> None:
        if self.m_sortedRecords is None:
            raise RuntimeError("refit needs a build(..., keep_topology=True) first")
        nV, nT = _check_indexed(vertices, indices)
        if nT != self.m_nTriangles:
            raise ValueError("triangle count differs from the build")
        check(self.ctx.lib.mlbvh_refit_indexed(self.ctx.handle, _ptr(vertices), nV, _ptr(indices), nT,
                                               _ptr(self.m_sortedRecords), _ptr(self.m_deltas),
                                               _ptr(self.m_internals), _ptr(self.m_rootNode), _stream(stream)))

    def empty(self) -> bool:
        return self.m_nTriangles == 0


def _check_indexed(vertices: torch.Tensor, indices: torch.Tensor):
    if vertices.dtype != torch.float32 or vertices.dim() != 2 or vertices.shape[1] != 3:
        raise ValueError("vertices must be (V, 3) float32")
    if indices.dtype != torch.int32 or indices.dim() != 2 or indices.shape[1] != 3:
        raise ValueError("indices must be (T, 3) int32 (u32 vertex indices)")
    return vertices.shape[0], indices.shape[0]


class RayGenerator:
    """camera.h RayGenerator; lookat is evaluated by the library on the host."""

    def __init__(self):
        self.g = RayGen((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0))

    def lookat(self, eye, center, up, fovy: float, imageWidth: int, imageHeight: int) -> "RayGenerator":
        f3 = ctypes.c_float * 3
        check(_capi.load().mlbvh_raygen_lookat(ctypes.byref(self.g), f3(*eye), f3(*center), f3(*up), fovy,
                                               imageWidth, imageHeight))
        return self

    def as_tuple(self):
        return tuple(self.g.origin), tuple(self.g.right), tuple(self.g.up)


def trace(builder_or_nodes, triangles: torch.Tensor, rays: torch.Tensor, query: int = RAY_QUERY_CLOSEST,
          root: Optional[torch.Tensor] = None, counters: Optional[torch.Tensor] = None, stream=None,
          device: int = 0, walk: str = "stack") -> torch.Tensor:
    """Trace rays (n, 8) float32 [ox,oy,oz,tmax,dx,dy,dz,pad]; returns hits as an
    (n, 8) int32 tensor holding the 32-byte Hit records.  ``walk``: "stack"
    (mlbvh_trace: intersect_stackfree's hits, per-lane LDS stack),
    "stackfree" (intersect_stackfree's own walk) or "escape_link"
    (intersect_escape_link)."""
    fns = {"stack": "mlbvh_trace", "stackfree": "mlbvh_trace_stackfree", "escape_link": "mlbvh_trace_escape_link"}
    if walk not in fns:
        raise ValueError(f"walk must be one of {sorted(fns)}, not {walk!r}")
    if isinstance(builder_or_nodes, BVHGPUBuilder):
        nodes, root = builder_or_nodes.m_internals, builder_or_nodes.m_rootNode
    else:
        nodes = builder_or_nodes
    n = rays.shape[0]
    hits = torch.empty((n, 8), dtype=torch.int32, device=rays.device)
    ctx = context(device)
    fn = getattr(ctx.lib, fns[walk])
    check(fn(ctx.handle, _ptr(nodes), _ptr(triangles), _ptr(root), _ptr(rays), _ptr(hits), n, query, _ptr(counters),
             _stream(stream)))
    return hits


def trace_indexed(builder_or_nodes, vertices: torch.Tensor, indices: torch.Tensor, rays: torch.Tensor,
                  query: int = RAY_QUERY_CLOSEST, root: Optional[torch.Tensor] = None,
                  counters: Optional[torch.Tensor] = None, stream=None, device: int = 0) -> torch.Tensor:
    """trace() over an indexed mesh (mlbvh_trace_indexed)."""
    if isinstance(builder_or_nodes, BVHGPUBuilder):
        nodes, root = builder_or_nodes.m_internals, builder_or_nodes.m_rootNode
    else:
        nodes = builder_or_nodes
    nV, _ = _check_indexed(vertices, indices)
    n = rays.shape[0]
    hits = torch.empty((n, 8), dtype=torch.int32, device=rays.device)
    ctx = context(device)
    check(ctx.lib.mlbvh_trace_indexed(ctx.handle, _ptr(nodes), _ptr(vertices), nV, _ptr(indices), _ptr(root),
                                      _ptr(rays), _ptr(hits), n, query, _ptr(counters), _stream(stream)))
    return hits


class BatchedBVH:
    """Many meshes built in one pipeline (mlbvh_build_batched).  ``tri_offsets``
    is the (nMeshes+1,) prefix of triangle counts; mesh m's nodes are rows
    ``[off[m]-m, off[m+1]-m-1)`` of ``nodes`` and its root is ``roots[m]``."""

    def __init__(self, tri_offsets, device: int = 0):
        import numpy as np

        self.ctx = context(device)
        off = np.asarray(tri_offsets, dtype=np.uint32)
        if off.ndim != 1 or off.shape[0] < 2 or off[0] != 0 or np.any(np.diff(off.astype(np.int64)) < 2):
            raise ValueError("tri_offsets must start at 0 and give every mesh >= 2 triangles")
        self.h_off = off
        self.n_meshes = off.shape[0] - 1
        self.total = int(off[-1])
        self.max_tris = int(np.diff(off.astype(np.int64)).max())
        dev = f"cuda:{device}"
        self.d_off = torch.from_numpy(off.view(np.int32)).to(dev)
        self.nodes = torch.empty((self.total - self.n_meshes, 16), dtype=torch.int32, device=dev)
        self.roots = torch.empty(self.n_meshes, dtype=torch.int32, device=dev)

    def build(self, triangles: torch.Tensor, buildOption: int = 0, stream=None) -> None:
        check(self.ctx.lib.mlbvh_build_batched(self.ctx.handle, _ptr(triangles), _ptr(self.d_off), self.n_meshes,
                                               self.total, self.max_tris, buildOption, _ptr(self.nodes),
                                               _ptr(self.roots), _stream(stream)))

    def trace(self, triangles: torch.Tensor, rays: torch.Tensor, query: int = RAY_QUERY_CLOSEST,
              hits: Optional[torch.Tensor] = None, counters: Optional[torch.Tensor] = None, stream=None):
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 8), dtype=torch.int32, device=rays.device)
        check(self.ctx.lib.mlbvh_trace_batched(self.ctx.handle, _ptr(self.nodes), _ptr(triangles), _ptr(self.d_off),
                                               _ptr(self.roots), self.n_meshes, _ptr(rays), _ptr(hits), n, query,
                                               _ptr(counters), _stream(stream)))
        return hits

    def build_trace_per_mesh(self, triangles: torch.Tensor, rays: torch.Tensor, ray_offsets,
                             query: int = RAY_QUERY_CLOSEST, buildOption: int = 0,
                             hits: Optional[torch.Tensor] = None, stream=None):
        """The plain path: one build + one trace launch sequence per mesh."""
        import numpy as np

        roff = np.ascontiguousarray(ray_offsets, dtype=np.uint32)
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 8), dtype=torch.int32, device=rays.device)
        check(self.ctx.lib.mlbvh_build_trace_per_mesh(
            self.ctx.handle, _ptr(triangles), self.h_off.ctypes.data_as(ctypes.c_void_p), self.n_meshes, buildOption,
            _ptr(self.nodes), _ptr(self.roots), _ptr(rays), roff.ctypes.data_as(ctypes.c_void_p), _ptr(hits), query,
            _stream(stream)))
        return hits


def set_tuning(sort_config: int, blocks_per_cu: int, chunk: int, device: int = 0) -> None:
    """Launch shapes (see mlbvh_set_tuning); results do not depend on them."""
    ctx = context(device)
    check(ctx.lib.mlbvh_set_tuning(ctx.handle, sort_config, blocks_per_cu, chunk))


def render_normal(pixels: torch.Tensor, width: int, height: int, rayGenerator: RayGenerator, rootNode, internals,
                  triangles, stream=None, device: int = 0) -> None:
    ctx = context(device)
    check(ctx.lib.mlbvh_render_normal(ctx.handle, _ptr(pixels), width, height, ctypes.byref(rayGenerator.g),
                                      _ptr(rootNode), _ptr(internals), _ptr(triangles), _stream(stream)))


def render_ao(pixels: torch.Tensor, width: int, height: int, rayGenerator: RayGenerator, rootNode, internals,
              triangles, useSobol: int = 1, nSample: int = 16, radius: float = 1.0, seed: int = 12345,
              ao: Optional[torch.Tensor] = None, counters: Optional[torch.Tensor] = None, stream=None,
              device: int = 0) -> None:
    ctx = context(device)
    check(ctx.lib.mlbvh_render_ao(ctx.handle, _ptr(pixels), _ptr(ao), width, height, ctypes.byref(rayGenerator.g),
                                  _ptr(rootNode), _ptr(internals), _ptr(triangles), useSobol, nSample, radius,
                                  seed & 0xFFFFFFFF, _ptr(counters), _stream(stream)))


def render_pt(accumulators: torch.Tensor, width: int, height: int, rayGenerator: RayGenerator, rootNode, internals,
              triangles, attribs, useSobol: int = 1, maxSPP: int = 1024, maxDepth: int = 1024, stream=None,
              device: int = 0) -> None:
    ctx = context(device)
    check(ctx.lib.mlbvh_render_pt(ctx.handle, _ptr(accumulators), width, height, ctypes.byref(rayGenerator.g),
                                  _ptr(rootNode), _ptr(internals), _ptr(triangles), _ptr(attribs), useSobol, maxSPP,
                                  maxDepth, _stream(stream)))


def pack(pixels: torch.Tensor, accumulators: torch.Tensor, n: int, stream=None, device: int = 0) -> None:
    ctx = context(device)
    check(ctx.lib.mlbvh_pack(ctx.handle, _ptr(pixels), _ptr(accumulators), n, _stream(stream)))


def deg2rad(d: float) -> float:
    return d * math.pi / 180.0
