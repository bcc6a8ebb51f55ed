"""vx_albedo_backward_noalloc: the deterministic backward with its workspace in
the library's fixed per-device array (no hipMalloc), against vx_albedo_backward.

Called through the C ABI. Two calls in a row, the second on a smaller image, so
the second runs over what the first left in the shared workspace; stored and
re-traced hits; the size limit and the missing row table are refused before any
launch. Gradients must be bit-identical to the allocating entry point's.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

DET = 1  # VX_GRAD_DETERMINISTIC


def _sync():
    import torch
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def setup(vx, cave, gpu_ctx):
    import torch
    grid, _, built = cave
    scene = vx.Scene(gpu_ctx, built)
    V = scene.voxel_count()  # builds the row table
    rng = np.random.default_rng(31)
    light = rng.normal(size=3)
    light = (light / np.linalg.norm(light)).astype(np.float32)
    albedo = torch.from_numpy(rng.random((V, 3), dtype=np.float32)).cuda()
    views = vx.seeded_views(grid, 32, 2)
    return scene, V, light, albedo, views


def _backward(vx, fn, scene, cam, w, h, light, gimg, rows, shade, V):
    import torch
    grad = torch.zeros((V, 3), dtype=torch.float32, device="cuda")
    st = fn(scene._h, C.byref(cam), w, h, light.ctypes.data, gimg.data_ptr(),
            rows.data_ptr() if rows is not None else None, shade.data_ptr() if shade is not None else None,
            grad.data_ptr(), DET, torch.cuda.current_stream().cuda_stream)
    return st, grad


@pytest.mark.parametrize("stored", [True, False])
def test_noalloc_matches_allocating_backward_twice(vx, setup, stored):
    import torch
    scene, V, light, albedo, views = setup
    # the second call is smaller and runs over the first call's workspace contents
    for cam, (w, h) in zip(views, ((1280, 720), (640, 360))):
        img, rows, shade = scene.albedo_forward(cam, w, h, light, albedo, store_hits=stored)
        g = torch.Generator(device="cuda").manual_seed(w)
        gimg = torch.randn((h, w, 3), generator=g, device="cuda", dtype=torch.float32)
        st_ref, ref = _backward(vx, vx.LIB.vx_albedo_backward, scene, cam, w, h, light, gimg, rows, shade, V)
        st_na, got = _backward(vx, vx.LIB.vx_albedo_backward_noalloc, scene, cam, w, h, light, gimg, rows, shade, V)
        _sync()
        assert st_ref == 0 and st_na == 0, vx.LIB.vx_last_error().decode()
        assert int((ref != 0).any(1).sum()) > 1000  # many voxels got a gradient
        assert torch.equal(got.view(torch.int32), ref.view(torch.int32)), f"{w}x{h} stored={stored}"


def test_noalloc_refuses_what_it_cannot_hold(vx, setup, cave, gpu_ctx):
    import torch
    scene, V, light, albedo, views = setup
    cam = views[0]
    w, h = 2048, 1024  # 2^21 pixels: over VX_NOALLOC_MAX_PIXELS
    gimg = torch.zeros((h, w, 3), dtype=torch.float32, device="cuda")
    st, _ = _backward(vx, vx.LIB.vx_albedo_backward_noalloc, scene, cam, w, h, light, gimg, None, None, V)
    assert st == 1 and "pixels" in vx.LIB.vx_last_error().decode()
    # a scene whose voxel row table was never built
    fresh = vx.Scene(gpu_ctx, cave[2])
    gimg = torch.zeros((36, 64, 3), dtype=torch.float32, device="cuda")
    st, _ = _backward(vx, vx.LIB.vx_albedo_backward_noalloc, fresh, cam, 64, 36, light, gimg, None, None, V)
    assert st == 1 and "row table" in vx.LIB.vx_last_error().decode()
    fresh.close()
