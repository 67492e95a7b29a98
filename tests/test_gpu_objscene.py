"""The OBJ scene path on the GPU: a Cornell box written as OBJ + MTL, loaded
into Triangle / TriangleAttrib arrays (main_gpu.cpp:209-282), built on the
GPU (nodes bit-identical to the oracle), rendered with the normal kernel
(bit-identical image) and path-traced (accumulators and packed image
bit-identical to the oracle's at the same Sobol streams)."""
import math
import os

import numpy as np
import pytest
import torch

from minimumlbvh_amd import objscene

pytestmark = pytest.mark.gpu


def test_cornell_box_obj_build_normal_pt(gpu, orc, tmp_path):
    sc = objscene.load_obj(objscene.cornell_box_obj(str(tmp_path)))
    tris, attribs = sc.triangles, sc.attribs
    n = tris.shape[0]
    t = torch.from_numpy(tris).cuda()
    b = gpu.BVHGPUBuilder()
    b.build(t, n, 0, gpu.OnesweepSort())
    torch.cuda.synchronize()
    on, oroot, _, _ = orc.build(tris)
    assert int(b.m_rootNode.cpu().numpy().view(np.uint32)[0]) == oroot
    assert np.array_equal(b.m_internals.cpu().numpy().view(np.uint32), on)

    eye, ctr, up, fov = objscene.CAMERA_CORNELL
    w, h, spp = 256, 256, 16
    g = gpu.RayGenerator().lookat(eye, ctr, up, math.radians(fov), w, h)
    og = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    px = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    gpu.render_normal(px, w, h, g, b.m_rootNode, b.m_internals, t)
    torch.cuda.synchronize()
    ref = orc.render_normal(og, w, h, oroot, on, tris)
    assert (ref != 0).mean() > 0.5
    assert np.array_equal(px.cpu().numpy().view(np.uint32), ref)

    at = torch.from_numpy(attribs).cuda()
    gacc = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    acc = np.zeros((w * h, 4), np.float32)
    for _ in range(spp):
        gpu.render_pt(gacc, w, h, g, b.m_rootNode, b.m_internals, t, at, 1, 1024, 4)
        acc = orc.render_pt(acc, og, w, h, oroot, on, tris, attribs, 1, 1024, 4)
    torch.cuda.synchronize()
    ga = gacc.cpu().numpy()
    print(f"cornell box: mean {acc[:, :3].mean():.5f}, {(ga != acc).any(axis=1).sum()} pixels differ", flush=True)
    assert acc[:, :3].mean() > 0
    assert np.array_equal(ga, acc)
    px = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    gpu.pack(px, gacc, w * h)
    torch.cuda.synchronize()
    assert np.array_equal(px.cpu().numpy().view(np.uint32), orc.pack(acc))


FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "cornelbox", "cornelbox.obj.gz")


def test_reference_cornell_box_scene(gpu, orc):
    """The reference demo's own scene (bin/assets/cornelbox.obj, committed as
    a fixture: 17,716 triangles with the head scan) with the demo's camera
    (main_gpu.cpp:160-161) and its three modes (main_gpu.cpp:320-378): GPU
    build bit-identical to the oracle's; the normal image, the AO image
    (Sobol, 16 rays of length 1, main_gpu.cu:111-177) with its occlusion
    fractions, and 8 spp of path tracing at depth 4 with the packed image,
    all bit-identical to the oracle's."""
    sc = objscene.load_obj(FIXTURE)
    tris, attribs = sc.triangles, sc.attribs
    n = tris.shape[0]
    assert n == 17716
    print(f"reference cornell box: {n} triangles", flush=True)
    t = torch.from_numpy(tris).cuda()
    b = gpu.BVHGPUBuilder()
    b.build(t, n, 0, gpu.OnesweepSort())
    torch.cuda.synchronize()
    on, oroot, _, _ = orc.build(tris)
    assert int(b.m_rootNode.cpu().numpy().view(np.uint32)[0]) == oroot
    assert np.array_equal(b.m_internals.cpu().numpy().view(np.uint32), on)

    eye, ctr, up, fov = objscene.CAMERA_CORNELL
    w, h, spp = 320, 240, 8
    g = gpu.RayGenerator().lookat(eye, ctr, up, math.radians(fov), w, h)
    og = orc.raygen_lookat(eye, ctr, up, math.radians(fov), w, h)
    px = torch.zeros(w * h, dtype=torch.int32, device="cuda")
    gpu.render_normal(px, w, h, g, b.m_rootNode, b.m_internals, t)
    torch.cuda.synchronize()
    ref = orc.render_normal(og, w, h, oroot, on, tris)
    assert (ref != 0).mean() > 0.45  # the room fills about half of the 4:3 frame
    assert np.array_equal(px.cpu().numpy().view(np.uint32), ref)

    ao = torch.zeros(w * h, dtype=torch.float32, device="cuda")
    gpu.render_ao(px, w, h, g, b.m_rootNode, b.m_internals, t, 1, 16, 1.0, 12345, ao=ao)
    torch.cuda.synchronize()
    rpx, rao, _ = orc.render_ao(og, w, h, oroot, on, tris, 1, 16, 1.0, 12345, primary_stack=True)
    assert (rao > 0).mean() > 0.45
    assert np.array_equal(ao.cpu().numpy(), rao)
    assert np.array_equal(px.cpu().numpy().view(np.uint32), rpx)

    at = torch.from_numpy(attribs).cuda()
    gacc = torch.zeros((w * h, 4), dtype=torch.float32, device="cuda")
    acc = np.zeros((w * h, 4), np.float32)
    for _ in range(spp):
        gpu.render_pt(gacc, w, h, g, b.m_rootNode, b.m_internals, t, at, 1, 1024, 4)
        acc = orc.render_pt(acc, og, w, h, oroot, on, tris, attribs, 1, 1024, 4)
    gpu.pack(px, gacc, w * h)
    torch.cuda.synchronize()
    ga = gacc.cpu().numpy()
    print(f"reference cornell box PT: mean {acc[:, :3].mean():.5f}, {(ga != acc).any(axis=1).sum()} pixels differ",
          flush=True)
    assert (acc[:, 0] > 0).sum() > 1000
    assert np.array_equal(ga, acc)
    assert np.array_equal(px.cpu().numpy().view(np.uint32), orc.pack(acc))
