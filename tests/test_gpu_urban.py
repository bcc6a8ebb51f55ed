"""Urban box grid at full size: 1024^3 voxels (7 levels above 8^3 bricks, far
pointers), 2560x1440 primary rays plus one seeded cosine-weighted bounce ray per
primary hit. Kernel vs oracle bit for bit on both ray sets, plus size-independent
properties (hit voxels are solid in the dense grid, t is the voxel-entry distance)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import global_voxels

pytestmark = pytest.mark.gpu

W, H = 2560, 1440


@pytest.fixture(scope="module")
def urban(vx, gpu_ctx):
    grid, cam = vx.gen_urban(1024, 3, 1.0)
    built = vx.build_from_dense(grid, bc1=True)
    return grid, cam, built, vx.Scene(gpu_ctx, built)


def _check_hits(grid, built, rays, hits):
    g = global_voxels(built, hits)
    hit = g[:, 0] >= 0
    assert (grid[g[hit, 2], g[hit, 1], g[hit, 0]] != 0).all()
    # t is the entry of the hit voxel's AABB along the normalised ray (voxel units, world = [0,N]^3)
    d = rays["d"][hit].astype(np.float64)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = rays["o"][hit].astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t0 = (g[hit] - o) / d
        t1 = (g[hit] + 1.0 - o) / d
    tn = np.nanmax(np.where(np.isfinite(t0), np.minimum(t0, t1), -np.inf), axis=1)
    tn = np.maximum(tn, 0.0)
    t = hits["t"][hit].astype(np.float64)
    assert np.all(np.abs(t - tn) <= 1e-5 * np.maximum(1.0, np.abs(tn)) + 1e-3)


def test_urban_tree_uses_far_pointers(urban):
    built = urban[2]
    assert built.node_count > 0x7FFF and (built.nodes[:, 0] & 0x8000).any()
    assert built.config.user_max_levels == 10 and built.brick_count > 100000


def test_urban_primary_and_bounce_bit_exact(vx, orc, urban):
    import torch
    grid, cam, built, scene = urban
    primary = vx.camera_rays(cam, W, H, 3.0e38)
    hits = scene.trace(primary)
    img, _ = scene.render(cam, W, H, tmax=3.0e38)
    bounce = scene.bounce_rays(primary, hits, seed=1234)
    bounce_hits = scene.trace(bounce)
    torch.cuda.synchronize()
    p_rays = primary.cpu().numpy().view(orc.RAY_DTYPE)
    b_rays = bounce.cpu().numpy().view(orc.RAY_DTYPE)
    got_p = vx.hits_to_numpy(hits)
    got_b = vx.hits_to_numpy(bounce_hits)
    assert vx.hits_to_numpy(img).tobytes() == got_p.tobytes()  # vx_render traces the same rays
    oscene = orc.OracleScene.from_built(built)
    want_p, _ = orc.trace(oscene, p_rays, nthreads=16)
    assert got_p.tobytes() == want_p.tobytes()
    want_b, _ = orc.trace(oscene, b_rays, nthreads=16)
    bad = np.nonzero((got_b.view(np.uint32).reshape(-1, 8) != want_b.view(np.uint32).reshape(-1, 8)).any(1))[0]
    assert len(bad) == 0, f"{len(bad)} bounce records differ; first {bad[:1]}: {got_b[bad[0]]} vs {want_b[bad[0]]}"
    hit_p = (got_p["flags"] & 1) == 1
    assert hit_p.mean() > 0.5
    # a bounce ray exists exactly for each primary hit and leaves the surface
    assert np.array_equal(np.any(b_rays["d"] != 0, axis=1), hit_p)
    _check_hits(grid, built, p_rays, got_p)
    _check_hits(grid, built, b_rays, got_b)
