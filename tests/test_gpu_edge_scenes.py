"""GPU parity on degenerate scenes, ragged launch sizes and tiny images.

Scenes: no bricks at all, every voxel solid (16^3 and 64^3), and a single voxel
in the far corner brick of a 16^3 grid. Ray batches of 0, 1, 255, 257 and
10 000 rays (partial waves and workgroups); images of 1x1, 7x3, 33x17 and
100x1 pixels (partial 16x16 tiles), from cameras outside the volume and inside
a solid voxel. Every hit record must equal the oracle's byte for byte, and the
entry-face RGBA8 image the oracle's render word for word.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import record_diff, row_rays

pytestmark = pytest.mark.gpu


def _grids():
    rng = np.random.default_rng(5)
    full64 = (rng.integers(0, 2 ** 24, (64,) * 3, dtype=np.uint32) | np.uint32(0xFF000000)).astype(np.uint32)
    corner = np.zeros((16,) * 3, np.uint32)
    corner[15, 15, 15] = 0xFF3366CC  # grid[z, y, x]
    return {
        "empty16": np.zeros((16,) * 3, np.uint32),
        "full16": np.full((16,) * 3, 0xFF8040C0, np.uint32),
        "full64": full64,
        "corner16": corner,
    }


@pytest.fixture(scope="module")
def edge_scenes(vx):
    return {name: (g, vx.build_from_dense(g, bc1=(name != "full16"))) for name, g in _grids().items()}


def _rays(orc, rng, n, N):
    r = np.zeros(n, orc.RAY_DTYPE)
    k = n // 2
    r["o"][:k] = rng.uniform(0, N, (k, 3))
    r["o"][k:] = rng.uniform(-0.5 * N, 1.5 * N, (n - k, 3))
    r["d"] = rng.normal(size=(n, 3))
    r["d"][: n // 10] = (1.0, 0.0, 0.0)  # axis-parallel
    r["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N * 0.3, n), 0.0)
    r["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N, n), 1e30)
    return r


def _check(got, want, label):
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{label}: {len(bad)} of {len(want)} records differ; first at {bad[0]}: got {got[bad[0]]} want {want[bad[0]]}"


@pytest.mark.parametrize("name", ["empty16", "full16", "full64", "corner16"])
def test_trace_ragged_batches_bit_exact(vx, orc, edge_scenes, gpu_ctx, name):
    import torch
    grid, built = edge_scenes[name]
    scene = vx.Scene(gpu_ctx, built)
    oscene = orc.OracleScene.from_built(built)
    rng = np.random.default_rng(11)
    for n in (1, 255, 257, 10_000):
        rays = _rays(orc, rng, n, built.N)
        got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
        torch.cuda.synchronize()
        want, _ = orc.trace(oscene, rays, nthreads=8)
        _check(got, want, f"{name} n={n}")
        if name == "empty16":
            assert not (got["flags"] & 1).any()
        if name == "full16":
            inside = (rays["o"] > 0).all(1) & (rays["o"] < built.N).all(1) & (rays["tmin"] == 0)
            valid = np.isfinite(rays["d"]).all(1)
            assert (got["flags"][inside & valid] & 1).all()  # starts in a solid voxel: always a hit
    # an empty batch is a no-op
    assert vx.LIB.vx_trace(scene._h, None, 0, None, None) == 0


@pytest.mark.parametrize("name", ["empty16", "full16", "full64", "corner16"])
def test_render_tiny_images_bit_exact(vx, orc, edge_scenes, gpu_ctx, name):
    import torch
    grid, built = edge_scenes[name]
    N = built.N
    scene = vx.Scene(gpu_ctx, built)
    oscene = orc.OracleScene.from_built(built)
    cams = [vx.make_camera((-0.3 * N, -0.2 * N, -0.5 * N), (1.0, 0.8, 1.2), 60.0, 1.5),
            vx.make_camera((0.5 * N + 0.25, 0.5 * N + 0.25, 0.5 * N + 0.25), (0.3, -1.0, 0.2), 90.0, 1.0)]
    for ci, cam in enumerate(cams):
        for w, h in ((1, 1), (7, 3), (33, 17), (100, 1)):
            hits, rgba = scene.render(cam, w, h, tmax=1e30, rgba=True)
            torch.cuda.synchronize()
            got = vx.hits_to_numpy(hits)
            rays = orc.camera_rays(cam.as_floats(), w, h, 1e30)
            want, _ = orc.trace(oscene, rays, nthreads=4)
            _check(got, want, f"{name} camera {ci} {w}x{h}")
            img = rgba.cpu().numpy().view(np.uint32)
            want_img = orc.render(oscene, cam.as_floats(), w, h, 1e30, nthreads=4)
            assert np.array_equal(img, want_img), f"{name} camera {ci} {w}x{h}: RGBA8 differs"


@pytest.mark.parametrize("N", [2048, 4096])
def test_maximum_size_sparse_world_bit_exact(vx, orc, gpu_ctx, N):
    """The deepest trees the traversal stack supports (8 and 9 levels above the
    bricks), built from a brick list: aimed and random rays, a camera frame, and
    the same with the top descriptor levels staged in LDS."""
    import torch
    from conftest import aimed_rays, sparse_world
    cells, voxels, solid = sparse_world(N, 21, n_scattered=40, cluster=3)
    built = vx.build_from_bricks(cells, voxels, N, bc1=True)
    scene = vx.Scene(gpu_ctx, built)
    oscene = orc.OracleScene.from_built(built)
    rng = np.random.default_rng(N + 1)
    rays = np.concatenate([aimed_rays(orc.RAY_DTYPE, rng, 20_000, N, solid), _rays(orc, rng, 20_000, N)])
    want, _ = orc.trace(oscene, rays, nthreads=8)
    assert (want["flags"] & 1).mean() > 0.4
    target = solid[len(solid) // 2] + 0.5
    eye = target + np.array([-30.0, 25.0, -35.0])
    cam = vx.make_camera(eye, target - eye, 60.0, 2.0)
    crays = orc.camera_rays(cam.as_floats(), 258, 129, 1e30)
    cwant, _ = orc.trace(oscene, crays, nthreads=8)
    assert (cwant["flags"] & 1).mean() > 0.05
    for depth in (0, 3):
        scene.set_lds_depth(depth)
        got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
        torch.cuda.synchronize()
        _check(got, want, f"N={N} rays, lds depth {depth}")
        hits, _ = scene.render(cam, 258, 129, tmax=1e30)
        torch.cuda.synchronize()
        _check(vx.hits_to_numpy(hits), cwant, f"N={N} camera, lds depth {depth}")


def test_scene_and_batch_outlive_their_context(vx, orc, edge_scenes):
    """Scenes and batches copy their context's device: destroying the context
    first leaves them usable and destroyable."""
    import torch
    grid, built = edge_scenes["full64"]
    ctx = vx.Context(0)
    scene = vx.Scene(ctx, built)
    batch = vx.Batch(ctx, [scene])
    ctx.close()
    rays = _rays(orc, np.random.default_rng(3), 4096, built.N)
    d_rays = vx.rays_to_device(rays)
    got = vx.hits_to_numpy(scene.trace(d_rays))
    hits = batch.trace(d_rays, np.array([0, len(rays)], np.uint64))
    torch.cuda.synchronize()
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    _check(got, want, "scene after its context")
    _check(vx.hits_to_numpy(hits), want, "batch after its context")
    batch.close()
    scene.close()


def test_iteration_cap_bit_exact(vx, orc, brick_row, gpu_ctx):
    """Rays that end at the 512-iteration cap, and rays just under it, through each form of vx_trace: a
    few rays (one launch, leaf visits postponed), 64 rays sharing one origin (a lockstep block) and 1.2 M
    rays (the two-launch form, the cap rays among lane-refilled incoherent ones); records bit-exact vs
    the oracle, the capped ones misses with iteration count 512."""
    N, built = brick_row
    scene = vx.Scene(gpu_ctx, built)
    oscene = orc.OracleScene.from_built(built)
    rng = np.random.default_rng(12)
    few = row_rays(orc.RAY_DTYPE, N, np.array([-1.0, 100.0, N - 2000.0, N - 200.0, N - 3000.0], np.float32))
    same = row_rays(orc.RAY_DTYPE, N, np.full(64, -1.0, np.float32))
    same["d"][:, 1:] = rng.uniform(-1e-4, 1e-4, (64, 2))  # stays in the row for most of it
    big = np.zeros(1_200_000, orc.RAY_DTYPE)
    big["o"] = rng.uniform(0, N, (len(big), 3))
    big["o"][:, 1:] = rng.uniform(0, 8, (len(big), 2))
    big["d"] = rng.normal(size=(len(big), 3))
    big["tmax"] = 1e30
    k = rng.choice(len(big), 4000, replace=False)
    big[k] = row_rays(orc.RAY_DTYPE, N, rng.uniform(-1, N - 100, len(k)).astype(np.float32))
    for label, rays in (("few", few), ("lockstep", same), ("two-launch", big)):
        want, _ = orc.trace(oscene, rays, nthreads=8)
        got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
        _check(got, want, label)
        assert ((want["flags"] >> 16) == 512).any(), label
    scene.close()
