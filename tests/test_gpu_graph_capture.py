"""Per-frame launches captured in a HIP graph.

vx_camera_rays, vx_render, vx_trace and vx_trace_lod only launch kernels on
the caller's stream: no allocation, no memset, no synchronisation. A caller may therefore
capture a frame (torch.cuda.CUDAGraph records through hipStreamBeginCapture)
and replay it. Every replay must reproduce the eager outputs bit for bit,
with the program's own eager work and synchronisations between replays.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

W, H = 320, 180


def test_captured_frame_replays_bit_exact(vx, cave, gpu_ctx):
    import torch
    grid, _, built = cave
    scene = vx.Scene(gpu_ctx, built)
    cam = vx.seeded_views(grid, 91, 1)[0]
    rays = vx.camera_rays(cam, W, H, 3.0e38)
    lod = vx.VxLod(0.5, 4.0 / grid.shape[0])

    ref_hits, ref_rgba = scene.render(cam, W, H, hits=True, rgba=True)
    ref_trace = scene.trace(rays)
    ref_lod = scene.trace_lod(rays, lod)
    torch.cuda.synchronize()
    ref = [t.cpu().numpy().copy() for t in (rays, ref_hits, ref_rgba, ref_trace, ref_lod)]
    assert (vx.hits_to_numpy(ref_trace)["flags"] & 1).any()

    rays_out = torch.empty_like(rays)
    trace_out = torch.empty_like(ref_trace)
    lod_out = torch.empty_like(ref_lod)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up off the default stream, as capture expects
        scene.render(cam, W, H, hits=True, rgba=True)
        scene.trace(rays, hits=trace_out)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()

    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        vx.camera_rays(cam, W, H, 3.0e38, out=rays_out)
        g_hits, g_rgba = scene.render(cam, W, H, hits=True, rgba=True)
        scene.trace(rays_out, hits=trace_out)
        scene.trace_lod(rays_out, lod, hits=lod_out)
    outs = (rays_out, g_hits, g_rgba, trace_out, lod_out)

    for replay in range(4):
        for t in outs:  # poison the outputs with eager work between replays
            t.fill_(0x5A if t.dtype == torch.uint8 else 0x5A5A5A5A)
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for name, got, want in zip(("camera rays", "render hits", "render rgba", "trace", "trace_lod"), outs, ref):
            assert np.array_equal(got.cpu().numpy(), want), f"replay {replay}: {name} differs from the eager launch"
    del graph
    scene.close()


def test_captured_frame_keeps_its_scene(vx, cave, gpu_ctx):
    # vx.CapturedFrame retains the scene: its creator closes it, the frame still replays bit-exact
    import torch
    grid, _, built = cave
    scene = vx.Scene(gpu_ctx, built)
    cam = vx.seeded_views(grid, 92, 1)[0]
    ref_hits, _ = scene.render(cam, W, H)
    torch.cuda.synchronize()
    ref = ref_hits.cpu().numpy().copy()
    out = torch.empty_like(ref_hits)

    def record():
        vx._check(vx.LIB.vx_render(scene._h, vx.C.byref(cam), W, H, 1e30, out.data_ptr(), None,
                                   torch.cuda.current_stream().cuda_stream), "vx_render")
        return out

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        record()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    frame = vx.CapturedFrame([scene], record, keep=(out,))
    scene.close()  # the frame's reference keeps the device buffers alive
    for _ in range(3):
        out.fill_(0x5A)
        torch.cuda.synchronize()
        frame.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), ref)
    frame.close()
