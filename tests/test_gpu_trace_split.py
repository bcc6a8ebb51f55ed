"""vx_trace's launches: a call whose rays give every resident refill wave (about
7000) at least 128 rays runs as two launches -- blocks of 64 rays that share
one origin traced in lockstep (k_trace), every other block with lane refill
over one equal range per wave of a scattered block order (k_trace_refill); a
shorter call is one k_trace launch that postpones leaf visits in incoherent
blocks. Every form must give the oracle's records whatever the mix.

The ray sets interleave camera blocks (one origin) with incoherent runs at
offsets that are not multiples of 64, so some blocks hold both kinds (they are
incoherent blocks), ranges cut through blocks, and counts are not multiples of
64. The large sizes reach past 128 rays per resident wave, so they take the
two-launch form, and waves take several blocks and several refills.

The first launch marks the parts of the blocks it leaves to the refill launch
in the hit buffer itself (the first record of each part); the output buffer's
earlier contents, marks included, must not matter.
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


@pytest.mark.parametrize("w,h,n_runs", [(333, 219, 0), (333, 219, 7), (333, 219, 400), (1440, 810, 0),
                                         (1440, 810, 9), (1440, 810, 3000)])
def test_mixed_coherent_and_incoherent_blocks(vx, orc, cave, gpu_ctx, w, h, n_runs):
    grid, cam, built = cave
    rng = np.random.default_rng(100 + n_runs)
    rays = _mixed_rays(orc, cam, w, h, 256, rng, n_runs)
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


@pytest.mark.parametrize("w,h,n_runs", [(320, 180, 50), (1440, 810, 2000)])
def test_mixed_blocks_lod(vx, orc, cave, gpu_ctx, w, h, n_runs):
    import math
    grid, cam, built = cave
    rng = np.random.default_rng(7)
    rays = _mixed_rays(orc, cam, w, h, 256, rng, n_runs)
    lod = vx.lod_with_bias(vx.lod_from_camera(math.radians(cam.fov_deg), h), 4.0)
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace_lod(vx.rays_to_device(rays), lod))
    _sync()
    want = orc.trace_lod(orc.OracleScene.from_built(built), rays, lod.ray_orig_size, lod.ray_dir_size, nthreads=8)
    assert ((got["flags"] >> 6) & 1).sum() > 0  # some LOD stops
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{len(bad)} of {len(rays)} LOD records differ, first at {bad[0]}"


def test_stale_output_contents_do_not_matter(vx, orc, cave, gpu_ctx):
    """The output buffer starts filled with the refill mark (flags 6: face-axis field 3) in every record, then
    with the records of another ray set; both calls must give the oracle's records."""
    import torch
    grid, cam, built = cave
    rng = np.random.default_rng(31)
    rays = _mixed_rays(orc, cam, 1440, 810, 256, rng, 2500)
    other = _mixed_rays(orc, cam, 1440, 810, 256, np.random.default_rng(32), 700)[: len(rays)]
    scene = vx.Scene(gpu_ctx, built)
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    hits = torch.zeros(len(rays) * 8, dtype=torch.int32, device="cuda")
    hits.view(-1, 8)[:, 0] = 6
    for prefill in ("marks", "other records"):
        if prefill == "other records":
            scene.trace(vx.rays_to_device(other), hits.view(torch.uint8))
        scene.trace(vx.rays_to_device(rays), hits.view(torch.uint8))
        _sync()
        got = vx.hits_to_numpy(hits.view(torch.uint8))
        bad = record_diff(got, want)
        assert len(bad) == 0, f"{prefill}: {len(bad)} of {len(rays)} records differ, first at {bad[0]}"
