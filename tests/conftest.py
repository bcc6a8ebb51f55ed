"""Shared fixtures: the package (vbvs--vixen_amd), the oracle and seeded scenes."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import __graft_entry__ as entry  # noqa: E402


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X); runs through the C ABI")


@pytest.fixture(scope="session")
def vx():
    return entry.load_package()


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle
    return oracle


@pytest.fixture(scope="session")
def cornell(vx):
    grid, cam = vx.gen_cornell(64, 1)
    return grid, cam, vx.build_from_dense(grid, bc1=False)


@pytest.fixture(scope="session")
def cornell_bc1(vx, cornell):
    grid, cam, _ = cornell
    return grid, cam, vx.build_from_dense(grid, bc1=True)


@pytest.fixture(scope="session")
def cave(vx):
    grid, cam = vx.gen_cave(256, 2, 0.5)
    return grid, cam, vx.build_from_dense(grid, bc1=True)


@pytest.fixture(scope="session")
def gpu_ctx(vx):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return vx.Context(0)


def global_voxels(built, hits: np.ndarray) -> np.ndarray:
    """Hit records -> global voxel coordinates (x, y, z), -1 for misses."""
    hit = (hits["flags"] & 1) == 1
    bc = built.brick_coords[np.where(hit, hits["brick"], 0)]
    lin = hits["voxel"].astype(np.int64)
    g = np.stack([bc[:, 0] * 8 + (lin & 7), bc[:, 1] * 8 + ((lin >> 3) & 7), bc[:, 2] * 8 + ((lin >> 6) & 7)], 1)
    return np.where(hit[:, None], g, -1)


def record_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Indices of hit records whose 32 bytes differ."""
    wa = a.view(np.uint32).reshape(-1, 8)
    wb = b.view(np.uint32).reshape(-1, 8)
    return np.nonzero((wa != wb).any(1))[0]


def leaving_rays(ray_dtype, N, n_per, seed):
    """Origins within 2^-24 N of a max face, heading out through it: the set-up
    succeeds but the normalised start rounds onto the face, so the ESVO interval
    is empty (a miss after 0 iterations, decided after the ray inputs are
    written)."""
    rng = np.random.default_rng(seed)
    near = np.float32(N) * np.float32(1 - 2 ** -24)
    rays = np.zeros(3 * n_per, ray_dtype)
    for a in range(3):
        sl = slice(a * n_per, (a + 1) * n_per)
        o = rng.uniform(1, N - 1, (n_per, 3)).astype(np.float32)
        o[:, a] = near
        d = rng.normal(size=(n_per, 3))
        d[:, a] = np.abs(d[:, a]) + 0.1
        rays["o"][sl], rays["d"][sl] = o, d
    rays["tmax"] = 3.0e38
    return rays


def sparse_world(N, seed, n_scattered=60, cluster=4, density=0.3):
    """A sparse world for vx_build_from_bricks: scattered bricks (two in opposite
    corner cells), a cluster of cluster^3 adjacent bricks, seeded RGBA8 voxels
    at the given density. Returns cells [n, 3] (x, y, z), voxels [n, 8, 8, 8]
    (z, y, x) and the global (x, y, z) of every solid voxel."""
    rng = np.random.default_rng(seed)
    B = N // 8
    cells = {(0, 0, 0), (B - 1, B - 1, B - 1)}
    while len(cells) < n_scattered:
        cells.add(tuple(int(v) for v in rng.integers(0, B, 3)))
    c0 = rng.integers(0, B - cluster, 3)
    for dz in range(cluster):
        for dy in range(cluster):
            for dx in range(cluster):
                cells.add((int(c0[0] + dx), int(c0[1] + dy), int(c0[2] + dz)))
    cells = np.array(sorted(cells), np.uint32)
    n = len(cells)
    solid = rng.random((n, 8, 8, 8)) < density
    solid[:, 0, 0, 0] = True  # no brick is empty
    vox = np.where(solid, rng.integers(0, 2 ** 24, (n, 8, 8, 8)).astype(np.uint32) | np.uint32(0xFF000000), 0)
    vox = vox.astype(np.uint32)
    b, z, y, x = np.nonzero(vox)
    solid_xyz = np.stack([cells[b, 0] * 8 + x, cells[b, 1] * 8 + y, cells[b, 2] * 8 + z], 1).astype(np.int64)
    perm = rng.permutation(n)
    return cells[perm], vox[perm], solid_xyz


def aimed_rays(ray_dtype, rng, n, N, solid_xyz):
    """Rays from random origins (half inside the volume, half around it) towards
    random points of random solid voxels."""
    r = np.zeros(n, ray_dtype)
    k = n // 2
    r["o"][:k] = rng.uniform(0, N, (k, 3))
    r["o"][k:] = rng.uniform(-0.2 * N, 1.2 * N, (n - k, 3))
    tgt = solid_xyz[rng.integers(0, len(solid_xyz), n)] + rng.uniform(0.05, 0.95, (n, 3))
    r["d"] = tgt - r["o"]
    r["tmax"] = 3.0e38
    return r


@pytest.fixture(scope="session")
def brick_row(vx):
    """A 4096^3 world holding one row of 512 bricks along x (y, z in [0, 8)). Bricks 0-510 hold one voxel
    at (0, 7, 7), beside the line y = z = 1.5; the last brick holds voxel (7, 1, 1), on it. A ray along
    +x on that line visits every brick it crosses, one ESVO iteration or more each, so from far enough
    back it reaches the reference's 512-iteration cap before the last brick."""
    N = 4096
    nb = N // 8
    cells = np.zeros((nb, 3), np.uint32)
    cells[:, 0] = np.arange(nb)
    vox = np.zeros((nb, 8, 8, 8), np.uint32)
    vox[:-1, 7, 7, 0] = 0xFF112233  # voxels[i, z, y, x]
    vox[-1, 1, 1, 7] = 0xFF445566
    return N, vx.build_from_bricks(cells, vox, N, bc1=False)


def row_rays(ray_dtype, N, starts):
    """Rays along +x on the line y = z = 1.5 from the given x."""
    r = np.zeros(len(starts), ray_dtype)
    r["o"][:, 0] = starts
    r["o"][:, 1:] = 1.5
    r["d"][:, 0] = 1.0
    r["tmax"] = 1e30
    return r
