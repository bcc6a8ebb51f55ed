"""vx_trace's two launches: blocks of 64 rays that share one origin are traced in
lockstep (k_trace), every other block with lane refill over one equal ray range
per wave (k_trace_refill). Both must give the oracle's records whatever the mix.

The ray sets interleave camera blocks (one origin) with incoherent runs at
offsets that are not multiples of 64, so some blocks hold both kinds (they are
incoherent blocks), ranges cut through blocks, and counts are not multiples of
64. Sizes reach past one range per resident wave (about 7000 waves), so waves
take several blocks and several refills.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import record_diff

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _mixed_rays(orc, cam, w, h, N, rng, n_runs):
    cam_rays = orc.camera_rays(cam.as_floats(), w, h, 1e30)
    parts, at = [], 0
    cuts = np.sort(rng.choice(np.arange(1, len(cam_rays)), n_runs, replace=False))
    for c in cuts:
        parts.append(cam_rays[at:c])
        k = int(rng.integers(1, 300))
        r = np.zeros(k, orc.RAY_DTYPE)
        r["o"] = rng.uniform(0, N, (k, 3))
        r["d"] = rng.normal(size=(k, 3))
        r["tmax"] = 3e38
        parts.append(r)
        at = c
    parts.append(cam_rays[at:])
    return np.concatenate(parts)


@pytest.mark.parametrize("n_runs", [0, 7, 400])
def test_mixed_coherent_and_incoherent_blocks(vx, orc, cave, gpu_ctx, n_runs):
    grid, cam, built = cave
    rng = np.random.default_rng(100 + n_runs)
    rays = _mixed_rays(orc, cam, 333, 219, 256, rng, n_runs)
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
    _sync()
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    bad = record_diff(got, want)
    assert len(bad) == 0, f"n_runs={n_runs}: {len(bad)} of {len(rays)} records differ, first at {bad[0]}"


@pytest.mark.parametrize("n", [1, 63, 64, 65, 4097, 1_000_003])
def test_incoherent_counts_and_ranges(vx, orc, cornell_bc1, gpu_ctx, n):
    # random rays only: every block goes to the refill launch; 1,000,003 rays give
    # every resident wave a range of over 128 rays that starts and ends inside blocks
    _, _, built = cornell_bc1
    rng = np.random.default_rng(n)
    r = np.zeros(n, orc.RAY_DTYPE)
    r["o"] = rng.uniform(-8, 72, (n, 3))
    r["d"] = rng.normal(size=(n, 3))
    r["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, 20, n), 0.0)
    r["tmax"] = 3e38
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(r)))
    _sync()
    want, _ = orc.trace(orc.OracleScene.from_built(built), r, nthreads=8)
    bad = record_diff(got, want)
    assert len(bad) == 0, f"n={n}: {len(bad)} records differ, first at {bad[0]}"


def test_mixed_blocks_lod(vx, orc, cave, gpu_ctx):
    import math
    grid, cam, built = cave
    rng = np.random.default_rng(7)
    rays = _mixed_rays(orc, cam, 320, 180, 256, rng, 50)
    lod = vx.lod_with_bias(vx.lod_from_camera(math.radians(cam.fov_deg), 180), 4.0)
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace_lod(vx.rays_to_device(rays), lod))
    _sync()
    want = orc.trace_lod(orc.OracleScene.from_built(built), rays, lod.ray_orig_size, lod.ray_dir_size, nthreads=8)
    assert ((got["flags"] >> 6) & 1).sum() > 0  # some LOD stops
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{len(bad)} of {len(rays)} LOD records differ, first at {bad[0]}"
