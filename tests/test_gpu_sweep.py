"""Seeded parity sweep: many small worlds, many kinds of rays, every entry point.

Each case draws a world (cave, urban, Cornell, a random brick list; edges 16 to
256; BC1 or RGBA8 colour; a random LDS staging depth) and a ray set mixing
random rays, rays aimed at solid voxels, rays from integer lattice points along
axis and diagonal directions (exact DDA and ESVO ties), clipped tmin/tmax
segments and invalid rays. vx_trace, vx_trace_lod (a random cone) and vx_render
(a small frame from a random camera) must all equal the oracle byte for byte.
The case label (index, world, edge, seed) is in every failure message.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import aimed_rays, record_diff, sparse_world

pytestmark = pytest.mark.gpu

CASES = 48


def _world(vx, rng, i):
    kind = ["cave", "urban", "cornell", "bricks"][i % 4]
    N = int(rng.choice([16, 32, 64, 128, 256])) if kind != "cornell" else int(rng.choice([32, 64]))
    seed = int(rng.integers(0, 2 ** 31))
    bc1 = bool(rng.integers(0, 2))
    if kind == "cave":
        grid, _ = vx.gen_cave(N, seed, float(rng.uniform(0.1, 0.9)))
        built = vx.build_from_dense(grid, bc1=bc1)
    elif kind == "urban":
        N = max(N, 64)
        grid, _ = vx.gen_urban(N, seed, float(rng.uniform(0.2, 1.0)))
        built = vx.build_from_dense(grid, bc1=bc1)
    elif kind == "cornell":
        grid, _ = vx.gen_cornell(N, seed)
        built = vx.build_from_dense(grid, bc1=bc1)
    else:
        B = N // 8
        cells, voxels, _ = sparse_world(N, seed, n_scattered=min(B ** 3, 20), cluster=min(B - 1, 2),
                                        density=float(rng.uniform(0.05, 0.6)))
        built = vx.build_from_bricks(cells, voxels, N, bc1=bc1)
    return kind, N, seed, built


def _solid_xyz(built):
    b, lin = np.nonzero(built.brick_data)
    c = built.brick_coords[b].astype(np.int64)
    return np.stack([c[:, 0] * 8 + (lin & 7), c[:, 1] * 8 + ((lin >> 3) & 7), c[:, 2] * 8 + (lin >> 6)], 1)


def _rays(orc, rng, N, solid, n=6000):
    parts = []
    r = np.zeros(n, orc.RAY_DTYPE)  # random, with clipped segments
    r["o"] = rng.uniform(-0.3 * N, 1.3 * N, (n, 3))
    r["d"] = rng.normal(size=(n, 3))
    r["tmin"] = np.where(rng.random(n) < 0.25, rng.uniform(0, N * 0.5, n), 0.0)
    r["tmax"] = np.where(rng.random(n) < 0.25, rng.uniform(0, N, n), 3e38)
    parts.append(r)
    if len(solid):
        parts.append(aimed_rays(orc.RAY_DTYPE, rng, n, N, solid))
    lat = np.zeros(n, orc.RAY_DTYPE)  # lattice origins, axis and diagonal directions: exact ties
    lat["o"] = rng.integers(0, N + 1, (n, 3)).astype(np.float32)
    dirs = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1], [1, 1, 1],
                     [2, 1, 0], [1, 2, 3]], np.float32)
    lat["d"] = dirs[rng.integers(0, len(dirs), n)] * rng.choice([-1.0, 1.0], (n, 3)).astype(np.float32)
    lat["tmax"] = 3e38
    parts.append(lat)
    bad = np.zeros(8, orc.RAY_DTYPE)  # invalid: zero, NaN and infinite directions or origins
    bad["o"] = [[1, 1, 1], [np.nan, 1, 1], [1, np.inf, 1], [2, 2, 2], [3, 3, 3], [1, 1, 1], [0, 0, 0], [5, 5, 5]]
    bad["d"] = [[0, 0, 0], [1, 0, 0], [0, 1, 0], [np.nan, 1, 1], [np.inf, 0, 0], [1e-30, 0, 0], [1, 1, 1], [1, 1, 1]]
    bad["tmax"] = [3e38, 3e38, 3e38, 3e38, 3e38, 3e38, -1.0, np.nan]
    parts.append(bad)
    return np.concatenate(parts)


@pytest.mark.parametrize("i", range(CASES))
def test_sweep_case_bit_exact(vx, orc, gpu_ctx, i):
    import torch
    rng = np.random.default_rng(1000 + i)
    kind, N, seed, built = _world(vx, rng, i)
    label = f"case {i}: {kind} N={N} seed={seed} bc1={built.bc1 is not None}"
    print(label, flush=True)
    scene = vx.Scene(gpu_ctx, built)
    scene.set_lds_depth(int(rng.integers(0, 4)))
    oscene = orc.OracleScene.from_built(built)
    rays = _rays(orc, rng, N, _solid_xyz(built))
    d_rays = vx.rays_to_device(rays)

    got = vx.hits_to_numpy(scene.trace(d_rays))
    torch.cuda.synchronize()
    want, _ = orc.trace(oscene, rays, nthreads=8)
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{label} vx_trace: {len(bad)} differ; first ray {rays[bad[0]]}"

    lod = vx.VxLod(float(rng.uniform(0, 2.0)), float(rng.uniform(0.01, 0.3)))
    got = vx.hits_to_numpy(scene.trace_lod(d_rays, lod))
    torch.cuda.synchronize()
    want = orc.trace_lod(oscene, rays, lod.ray_orig_size, lod.ray_dir_size, nthreads=8)
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{label} vx_trace_lod {lod.ray_orig_size, lod.ray_dir_size}: {len(bad)} differ; " \
                          f"first ray {rays[bad[0]]}"

    pos = rng.uniform(-0.2 * N, 1.2 * N, 3)
    cam = vx.make_camera(pos, rng.uniform(0.2 * N, 0.8 * N, 3) - pos, float(rng.uniform(30, 100)),
                         float(rng.uniform(0.5, 2.0)))
    w, h = int(rng.integers(1, 97)), int(rng.integers(1, 65))
    hits, rgba = scene.render(cam, w, h, tmax=1e30, rgba=True)
    torch.cuda.synchronize()
    crays = orc.camera_rays(cam.as_floats(), w, h, 1e30)
    want, _ = orc.trace(oscene, crays, nthreads=4)
    bad = record_diff(vx.hits_to_numpy(hits), want)
    assert len(bad) == 0, f"{label} vx_render {w}x{h}: {len(bad)} differ"
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), orc.render(oscene, cam.as_floats(), w, h, 1e30)), \
        f"{label} vx_render {w}x{h}: RGBA8 differs"
