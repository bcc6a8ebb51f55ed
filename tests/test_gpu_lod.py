"""GPU parity of vx_trace_lod (castRayWithLOD) against the oracle's orc_trace_lod.

Raw 32-byte hit records must be identical: LOD stops (flag, user scale,
iterations, parent descriptor, t bits) and every record that ran to full detail.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import record_diff

pytestmark = pytest.mark.gpu


def _same(got, want, label):
    bad = record_diff(got, want)
    assert len(bad) == 0, (f"{label}: {len(bad)} of {len(want)} records differ; first at {bad[0]}: "
                           f"got {got[bad[0]]} want {want[bad[0]]}")


def _gpu_lod(vx, scene, rays, lod):
    import torch
    hits = scene.trace_lod(vx.rays_to_device(rays), lod)
    torch.cuda.synchronize()
    return vx.hits_to_numpy(hits)


@pytest.mark.parametrize("lds_depth", [0, 3])
def test_camera_rays_lod_bit_exact(vx, orc, cave, gpu_ctx, lds_depth):
    _, cam, built = cave
    scene = vx.Scene(gpu_ctx, built)
    scene.set_lds_depth(lds_depth)
    osc = orc.OracleScene.from_built(built)
    W, H = 480, 270
    rays = orc.camera_rays(cam.as_floats(), W, H, 1e30)
    base = vx.lod_from_camera(math.radians(cam.fov_deg), H)
    fired = []
    for bias, near in ((0.0, None), (3.0, None), (5.0, None), (7.0, 2.0)):
        p = vx.lod_with_bias(vx.lod_from_camera(math.radians(cam.fov_deg), H, near) if near else base, bias)
        print(f"lod bias {bias} near {near} lds {lds_depth}", flush=True)
        got = _gpu_lod(vx, scene, rays, p)
        want = orc.trace_lod(osc, rays, p.ray_orig_size, p.ray_dir_size, nthreads=8)
        _same(got, want, f"bias {bias} near {near}")
        fired.append(float(((want["flags"] >> 6) & 1).mean()))
    assert fired[-1] > 0.3 and fired == sorted(fired)


def test_random_rays_lod_bit_exact(vx, orc, cave, cornell_bc1, gpu_ctx):
    rng = np.random.default_rng(13)
    for _, _, built in (cave, cornell_bc1):
        scene = vx.Scene(gpu_ctx, built)
        N = built.N
        n = 200_000
        rays = np.zeros(n, orc.RAY_DTYPE)
        k = n // 2
        rays["o"][:k] = rng.uniform(0, N, (k, 3))
        rays["o"][k:] = rng.uniform(-0.5 * N, 1.5 * N, (n - k, 3))
        rays["d"] = rng.normal(size=(n, 3))
        rays["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N * 0.3, n), 0.0)
        rays["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N, n), 1e30)
        lod = vx.VxLod(0.5, 8.0 / N)
        print(f"random lod rays N={N}", flush=True)
        got = _gpu_lod(vx, scene, rays, lod)
        want = orc.trace_lod(orc.OracleScene.from_built(built), rays, lod.ray_orig_size, lod.ray_dir_size,
                             nthreads=8)
        _same(got, want, f"random rays N={N}")
        assert ((want["flags"] >> 6) & 1).any()


def test_disabled_lod_is_vx_trace(vx, orc, cornell_bc1, gpu_ctx):
    import torch
    _, cam, built = cornell_bc1
    scene = vx.Scene(gpu_ctx, built)
    rays = vx.rays_to_device(orc.camera_rays(cam.as_floats(), 256, 256, 1e30))
    plain = vx.hits_to_numpy(scene.trace(rays))
    for dir_size in (0.0, -1.0, float("nan")):
        got = vx.hits_to_numpy(scene.trace_lod(rays, vx.VxLod(3.0, dir_size)))
        torch.cuda.synchronize()
        assert got.tobytes() == plain.tobytes()
