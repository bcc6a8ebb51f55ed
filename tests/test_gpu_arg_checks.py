"""The Python mirror refuses buffers the kernels cannot use, before any launch:
tensors on another device or not contiguous, hit buffers smaller than 32
bytes per ray, batched ray offsets of the wrong length, decreasing or past
the rays (the C side reads offsets[n_scenes] and trusts the rest), and
undersized pools of device-buffer scenes."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_undersized_and_foreign_buffers_are_refused(vx, cornell_bc1, gpu_ctx):
    import torch
    _, cam, built = cornell_bc1
    scene = vx.Scene(gpu_ctx, built)
    rays = vx.camera_rays(cam, 16, 16, 1e30)
    with pytest.raises(ValueError, match="bytes"):
        scene.trace(rays, hits=torch.empty(32 * 255, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError, match="bytes"):
        scene.trace_lod(rays, vx.VxLod(0.0, 0.01), hits=torch.empty(32, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError, match="bytes"):
        scene.trace_counters(rays, hits=torch.empty(8, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError, match="cuda:0"):
        scene.trace(rays.cpu())
    with pytest.raises(ValueError, match="contiguous"):
        scene.trace(torch.empty((512, 32), dtype=torch.uint8, device="cuda").t())
    with pytest.raises(ValueError, match="32-byte"):
        scene.trace(rays[:100])
    with pytest.raises(ValueError, match="bytes"):
        scene.bounce_rays(rays, torch.empty(64, dtype=torch.uint8, device="cuda"), seed=1)
    with pytest.raises(ValueError, match="bytes"):
        vx.camera_rays(cam, 16, 16, 1e30, out=torch.empty(100, dtype=torch.uint8, device="cuda"))
    # a correct call still works after the refusals
    good = scene.trace(rays, hits=torch.empty(32 * 256, dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()
    assert (vx.hits_to_numpy(good)["flags"] & 1).any()


def test_batched_offsets_and_hits_are_checked(vx, cornell_bc1, gpu_ctx):
    import torch
    _, cam, built = cornell_bc1
    s1, s2 = vx.Scene(gpu_ctx, built), vx.Scene(gpu_ctx, built)
    batch = vx.Batch(gpu_ctx, [s1, s2])
    rays = vx.camera_rays(cam, 16, 16, 1e30)
    hits = torch.empty_like(rays)
    for bad in ([0, 256], [0, 200, 100], [0, 100, 257]):
        with pytest.raises(ValueError):
            batch.trace(rays, np.array(bad, np.uint64), hits)
    with pytest.raises(ValueError, match="bytes"):
        batch.trace(rays, np.array([0, 100, 256], np.uint64), torch.empty(32 * 200, dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError):
        batch.graph(rays, np.array([0, 100], np.uint64), hits)
    out = batch.trace(rays, np.array([0, 100, 256], np.uint64), hits)
    torch.cuda.synchronize()
    assert (vx.hits_to_numpy(out)["flags"] & 1).any()


def test_device_scene_pools_are_checked(vx, cornell_bc1, gpu_ctx):
    import torch
    _, _, built = cornell_bc1
    t = {k: torch.from_numpy(np.ascontiguousarray(getattr(built, k)).view(np.uint8).reshape(-1)).cuda()
         for k in ("nodes", "brick_data", "bc1")}
    with pytest.raises(ValueError, match="bc1"):
        vx.Scene.from_device(gpu_ctx, t["nodes"], t["brick_data"], built.config, bc1=t["bc1"][:256])
    with pytest.raises(ValueError, match="cuda:0"):
        vx.Scene.from_device(gpu_ctx, t["nodes"].cpu(), t["brick_data"], built.config)
    sc = vx.Scene.from_device(gpu_ctx, t["nodes"], t["brick_data"], built.config, bc1=t["bc1"])
    sc.close()
