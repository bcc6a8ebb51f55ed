"""The Python mirror refuses buffers the kernels cannot use, before any launch:
tensors on another device or not contiguous, hit buffers smaller than 32
bytes per ray or overlapping the rays, batched ray offsets of the wrong length, decreasing or past
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
    with pytest.raises(ValueError, match="overlaps"):
        scene.trace(rays, hits=rays)                              # vx_trace reads rays after writing hits
    both = torch.empty(32 * 300, dtype=torch.uint8, device="cuda")
    both[: 32 * 256] = rays
    with pytest.raises(ValueError, match="overlaps"):
        scene.trace_lod(both[: 32 * 256], vx.VxLod(0.0, 0.01), hits=both[32 * 40:])
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


def test_device_scene_changed_in_place_is_refused_until_revalidated(vx, orc, cornell, gpu_ctx):
    # RGBA8 scene (no BC1) over device tensors: an in-place change blocks launches (torch's version
    # counter); revalidate() re-checks the tree, rebuilds the occupancy tables and the trace then matches the
    # oracle on the changed buffers. A broken tree is refused by the C ABI itself, with no launch.
    import ctypes as C
    import torch
    from conftest import record_diff
    grid, cam, built = cornell
    nodes = torch.from_numpy(np.ascontiguousarray(built.nodes).view(np.uint8).reshape(-1).copy()).cuda()
    bricks = torch.from_numpy(np.ascontiguousarray(built.brick_data).reshape(-1).view(np.int32).copy()).cuda()
    sc = vx.Scene.from_device(gpu_ctx, nodes, bricks, built.config)
    rays = orc.camera_rays(cam.as_floats(), 64, 64, 1e4)
    d_rays = vx.rays_to_device(rays)
    sc.trace(d_rays)
    # empty every other brick in place
    view = bricks.view(-1, 512)
    view[::2] = 0
    with pytest.raises(vx.VixenError, match="revalidate"):
        sc.trace(d_rays)
    sc.revalidate()
    got = vx.hits_to_numpy(sc.trace(d_rays))
    torch.cuda.synchronize()
    changed = built.brick_data.copy()
    changed[::2] = 0
    b2 = vx.BuiltScene(built.nodes, changed, None, built.brick_coords, built.internal_count, built.config)
    want, _ = orc.trace(orc.OracleScene.from_built(b2), rays, nthreads=8)
    assert len(record_diff(got, want)) == 0
    # a child pointer past the descriptor array: revalidation fails and the scene refuses launches
    bad = built.nodes.copy()
    root = int(bad[0, 0])
    bad[0, 0] = (root & ~0x7FFF) | 0x7FFF if not root & 0x8000 else root
    if root & 0x8000:
        bad[0, 1] = 0xFFFFFF
    nodes.copy_(torch.from_numpy(bad.view(np.uint8).reshape(-1).copy()).cuda())
    with pytest.raises(vx.VixenError, match="VX_ERR_INVALID_SCENE"):
        sc.revalidate()
    hits = torch.empty_like(d_rays)
    st = vx.LIB.vx_trace(sc._h, d_rays.data_ptr(), len(rays), hits.data_ptr(), None)
    assert st == 2 and b"revalidate" in vx.LIB.vx_last_error()
    # restored: accepted again
    nodes.copy_(torch.from_numpy(np.ascontiguousarray(built.nodes).view(np.uint8).reshape(-1).copy()).cuda())
    sc.revalidate()
    got = vx.hits_to_numpy(sc.trace(d_rays))
    torch.cuda.synchronize()
    assert len(record_diff(got, want)) == 0
    sc.close()
