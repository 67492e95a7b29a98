"""Differentiable primary-hit rendering on the GPU BVH.

``PrimaryRenderer(faces, width, height, camera, light)`` renders an indexed
fp32 mesh to a (height, width, 2) image of (hit distance t, max(n.l, 0)).
Its ``render(vertices)`` is differentiable in the vertices through
``PrimaryRenderFn`` (a torch.autograd.Function): the forward gathers the
triangles, refits the BVH (a full rebuild every ``rebuild_every`` calls) and
traces the primary rays; the backward runs the hand-written adjoint kernel
(mlbvh_primary_adjoint), which scatters dL/dvertex with fp32 atomics, holding
each pixel's hit triangle fixed.
"""
from __future__ import annotations

import ctypes

import torch

from . import lbvh
from ._capi import check


class PrimaryRenderer:
    def __init__(self, faces: torch.Tensor, width: int, height: int, camera: "lbvh.RayGenerator", light,
                 rebuild_every: int = 20, device: int = 0):
        if faces.dtype != torch.int32 or faces.dim() != 2 or faces.shape[1] != 3:
            raise ValueError("faces must be an (F, 3) int32 tensor of vertex indices")
        self.ctx = lbvh.context(device)
        self.faces = faces.contiguous()
        self.n_tris = faces.shape[0]
        self.width, self.height = width, height
        self.camera = camera
        self.light = (ctypes.c_float * 3)(*[float(x) for x in light])
        self.rebuild_every = rebuild_every
        self.builder = lbvh.BVHGPUBuilder(device=device)
        self.tris = torch.empty((self.n_tris, 9), dtype=torch.float32, device=faces.device)
        self.calls = 0
        self.hit_tri = torch.empty(width * height, dtype=torch.int32, device=faces.device)

    def update_bvh(self, vertices: torch.Tensor) -> None:
        check(self.ctx.lib.mlbvh_gather_triangles(self.ctx.handle, lbvh._ptr(vertices), lbvh._ptr(self.faces),
                                                   self.n_tris, lbvh._ptr(self.tris), lbvh._stream(None)))
        if self.calls % self.rebuild_every == 0 or self.builder.m_sortedRecords is None:
            self.builder.build(self.tris, self.n_tris, 0, keep_topology=True)
        else:
            self.builder.refit(self.tris)
        self.calls += 1

    def forward_raw(self, vertices: torch.Tensor) -> torch.Tensor:
        self.update_bvh(vertices)
        out = torch.empty((self.height, self.width, 2), dtype=torch.float32, device=vertices.device)
        check(self.ctx.lib.mlbvh_render_primary(self.ctx.handle, lbvh._ptr(out), lbvh._ptr(self.hit_tri), self.width,
                                                 self.height, ctypes.byref(self.camera.g), self.light,
                                                 lbvh._ptr(self.builder.m_rootNode),
                                                 lbvh._ptr(self.builder.m_internals), lbvh._ptr(self.tris),
                                                 lbvh._stream(None)))
        return out

    def adjoint(self, grad_out: torch.Tensor, vertices: torch.Tensor, hit_tri: torch.Tensor) -> torch.Tensor:
        grad = torch.zeros_like(vertices)
        check(self.ctx.lib.mlbvh_primary_adjoint(self.ctx.handle, lbvh._ptr(grad_out.contiguous()),
                                                  lbvh._ptr(hit_tri), self.width, self.height,
                                                  ctypes.byref(self.camera.g), self.light, lbvh._ptr(vertices),
                                                  lbvh._ptr(self.faces), lbvh._ptr(grad), lbvh._stream(None)))
        return grad

    def render(self, vertices: torch.Tensor) -> torch.Tensor:
        return PrimaryRenderFn.apply(vertices, self)


class PrimaryRenderFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vertices: torch.Tensor, renderer: PrimaryRenderer):
        v = vertices.detach().contiguous()
        out = renderer.forward_raw(v)
        ctx.renderer = renderer
        ctx.hit_tri = renderer.hit_tri.clone()
        ctx.save_for_backward(v)
        return out

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (v,) = ctx.saved_tensors
        return ctx.renderer.adjoint(grad_out, v, ctx.hit_tri), None
