"""The deterministic albedo reduction's fixed summation order, pinned bit for bit.

vx_albedo_backward in deterministic mode sorts (row, pixel) pairs by row (stable:
pixel order within a row) and sums each row's contributions dL/dpixel * shade in
a fixed order:
  * a row hit by at most 32 pixels: sequentially, in pixel order;
  * a row hit by 33 to 4096 pixels: 64 lanes, lane t adding the row's elements
    t, t+128, ... into one partial sum and t+64, t+192, ... into another, then
    s0 + s1 per lane and a halving tree over the lanes (h = 32, ..., 1);
  * a longer row: the same with 256 threads (tree h = 128, ..., 1);
and adds each row's sum to grad. With stored hits the per-pixel rows and shades
are inputs, so adversarial layouts are fed directly: rows of exactly 32 and 33
pixels (the short/long boundary), of 4096 and 4097 (the wave/workgroup
boundary), rows that cross the reduction's 512-position windows, one row
covering 100 000 pixels (its end lies windows past its start), rows 0 and V-1,
misses. The expected gradient is that order restated with
numpy float32 arithmetic; the GPU result must equal it bit for bit, from a
nonzero starting gradient (the backward accumulates). The cornell scene's row
keys need two 8-bit sort passes, the cave's three.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from albedo_order import MISS, adversarial_rows, restated_gradient

pytestmark = pytest.mark.gpu


def _backward(vx, fn, scene, cam, w, h, light, gimg, rows, shade, grad):
    import torch
    st = fn(scene._h, C.byref(cam), w, h, light.ctypes.data, gimg.data_ptr(), rows.data_ptr(), shade.data_ptr(),
            grad.data_ptr(), 1, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    vx._check(st, fn.__name__)


@pytest.mark.parametrize("scene_name", ["cornell", "cave"])
def test_fixed_order_adversarial_rows(vx, request, gpu_ctx, scene_name):
    import torch
    grid, cam, built = request.getfixturevalue(scene_name)
    scene = vx.Scene(gpu_ctx, built)
    V = scene.voxel_count()
    rng = np.random.default_rng(2024)
    w, h = 1024, 512
    rows = adversarial_rows(rng, w * h, V)
    shade = rng.random(w * h, dtype=np.float32)
    gimg = rng.standard_normal((h, w, 3), dtype=np.float32)
    grad0 = rng.standard_normal((V, 3), dtype=np.float32)
    light = np.array([0.0, 1.0, 0.0], np.float32)
    want = restated_gradient(rows, shade, gimg, grad0)
    d_rows = torch.from_numpy(rows.view(np.int32)).cuda()
    d_shade, d_gimg = torch.from_numpy(shade).cuda(), torch.from_numpy(gimg).cuda()
    for fn in (vx.LIB.vx_albedo_backward, vx.LIB.vx_albedo_backward_noalloc):
        grad = torch.from_numpy(grad0).cuda()
        _backward(vx, fn, scene, cam, w, h, light, d_gimg, d_rows, d_shade, grad)
        got = grad.cpu().numpy()
        bad = got.view(np.uint32) != want.view(np.uint32)
        assert not bad.any(), (f"{fn.__name__}: {int(bad.sum())} of {bad.size} gradient entries differ, "
                               f"first at row {int(np.argwhere(bad)[0][0])}")
    scene.close()


@pytest.mark.parametrize("case", ["one_pixel", "all_miss", "one_row_everywhere", "ragged"])
def test_fixed_order_small_and_degenerate(vx, cornell, gpu_ctx, case):
    import torch
    grid, cam, built = cornell
    scene = vx.Scene(gpu_ctx, built)
    V = scene.voxel_count()
    rng = np.random.default_rng(7)
    w, h = {"one_pixel": (1, 1), "all_miss": (37, 3), "one_row_everywhere": (700, 300), "ragged": (513, 7)}[case]
    px = w * h
    if case == "all_miss":
        rows = np.full(px, MISS, np.uint32)
    elif case == "one_row_everywhere":
        rows = np.full(px, 5, np.uint32)  # a single run ending at the last position
    else:
        rows = rng.integers(0, min(V, 40), px).astype(np.uint32)
    shade = rng.random(px, dtype=np.float32)
    gimg = rng.standard_normal((h, w, 3), dtype=np.float32)
    grad0 = rng.standard_normal((V, 3), dtype=np.float32)
    want = restated_gradient(rows, shade, gimg, grad0)
    grad = torch.from_numpy(grad0).cuda()
    _backward(vx, vx.LIB.vx_albedo_backward, scene, cam, w, h, np.array([0.0, 1.0, 0.0], np.float32),
              torch.from_numpy(gimg).cuda(), torch.from_numpy(rows.view(np.int32)).cuda(),
              torch.from_numpy(shade).cuda(), grad)
    got = grad.cpu().numpy()
    assert got.tobytes() == want.tobytes()
    scene.close()


def test_stored_rows_out_of_range_count_as_misses(vx, cornell, gpu_ctx):
    """Stored rows at or above the voxel count (a corrupt or foreign hit_row buffer) count as misses in both
    reductions: no gradient index leaves the array. Deterministic: bit-exact against the restated order with
    those pixels as misses. Atomic: within float32 rounding of it, and rows no valid pixel hit keep their
    starting values exactly."""
    import torch
    grid, cam, built = cornell
    scene = vx.Scene(gpu_ctx, built)
    V = scene.voxel_count()
    rng = np.random.default_rng(9)
    w, h = 320, 200
    px = w * h
    rows = rng.integers(0, min(V, 3000), px).astype(np.uint32)
    bad = rng.random(px) < 0.2
    rows[bad] = rng.choice(np.array([V, V + 1, 2 ** 31, 0xFFFFFFFE], np.uint32), int(bad.sum()))
    shade = rng.random(px, dtype=np.float32)
    gimg = rng.standard_normal((h, w, 3), dtype=np.float32)
    grad0 = rng.standard_normal((V, 3), dtype=np.float32)
    want = restated_gradient(np.where(bad, np.uint32(MISS), rows), shade, gimg, grad0)
    d_rows = torch.from_numpy(rows.view(np.int32)).cuda()
    d_shade, d_gimg = torch.from_numpy(shade).cuda(), torch.from_numpy(gimg).cuda()
    light = np.array([0.0, 1.0, 0.0], np.float32)
    for mode in (1, 0):
        grad = torch.from_numpy(grad0).cuda()
        st = vx.LIB.vx_albedo_backward(scene._h, C.byref(cam), w, h, light.ctypes.data, d_gimg.data_ptr(),
                                       d_rows.data_ptr(), d_shade.data_ptr(), grad.data_ptr(), mode,
                                       torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        vx._check(st, "vx_albedo_backward")
        got = grad.cpu().numpy()
        if mode == 1:
            assert got.tobytes() == want.tobytes()
        else:
            untouched = np.ones(V, bool)
            untouched[rows[~bad]] = False
            assert got[untouched].tobytes() == grad0[untouched].tobytes()
            assert np.allclose(got, want, rtol=1e-5, atol=1e-5)
    scene.close()
