"""The C++ ISVOStructure-shaped adapter (include/vixen_svo.hpp) on the host:
its structure queries and its record -> RayHit conversion, driven from the C++
host tests/svo_consumer (no GPU involved).

voxelExists / getChildMask (LaineKarrasOctree.cpp:108-170, 259-314) descend from
the root by the position's octant at each level; restated on the dense grid they
are: voxelExists(p, s) holds when p is in the world box and the cell of depth
min(s, D) holding p has a solid voxel (D = the brick level, where the descent
meets a leaf child); getChildMask(p, s) is, for s < D and that cell populated,
the bit set of its populated children (bit x | 2y | 4z), else 0. The positions
include cell corners, where the >= tests of the descent decide the octant.
"""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import svo_expect as se


def _run(*args, timeout=120):
    if not se.EXE.exists():
        pytest.fail("tests/svo_consumer is not built (run __graft_entry__.build())")
    r = subprocess.run([str(se.EXE), *map(str, args)], capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, r.stdout + r.stderr


def _cells_populated(occ: np.ndarray, depth: int) -> np.ndarray:
    """[z, y, x] populated flags of the 2^depth cells per axis."""
    N = occ.shape[0]
    k = 1 << depth
    c = N // k
    return occ.reshape(k, c, k, c, k, c).any(axis=(1, 3, 5))


def test_structure_queries(vx, tmp_path):
    out = tmp_path / "q.bin"
    _run("queries", out)
    raw = out.read_bytes()
    at = 0

    def take(dt, n=1):
        nonlocal at
        a = np.frombuffer(raw, dt, n, at)
        at += a.nbytes
        return a

    grid, _ = vx.gen_cave(64, 3, 0.3)
    built = vx.build_from_dense(grid, bc1=False)
    occ = grid != 0
    max_levels = int(take("<i4")[0])
    voxel_count, mem = (int(x) for x in take("<u8", 2))
    wmin, wmax = take("<f4", 3), take("<f4", 3)
    sizes = take("<f4", 13)
    bounds = take("<f4", 6)
    n_hier, n_attr, n_col = (int(x) for x in take("<u8", 3))
    assert max_levels == 6
    assert voxel_count == int(occ.sum())
    nb = len(built.brick_data) // 512 if built.brick_data.ndim == 1 else built.brick_data.shape[0]
    assert mem == built.nodes.nbytes + built.brick_data.nbytes + nb * 68
    assert list(wmin) == [0, 0, 0] and list(wmax) == [64, 64, 64]
    # getVoxelSize: x extent / 2^scale below getMaxLevels, 0 from there on
    assert list(sizes) == [64.0 / 2 ** s if s < 6 else 0.0 for s in range(13)]
    assert list(bounds) == [0, 0, 0, 64, 64, 64]
    assert (n_hier, n_attr, n_col) == (built.nodes.nbytes, built.brick_data.nbytes, 0)

    n = int(take("<u8")[0])
    dt = np.dtype([("p", "<f4", 3), ("q", "u1", (9, 2))])
    rows = take(dt, n)
    stats = raw[at + 8:].decode()
    assert at + 8 + int(np.frombuffer(raw, "<u8", 1, at)[0]) == len(raw)

    D = 3  # 64 / 8 = 8 bricks per axis
    p = rows["p"].astype(np.float64)
    inside = ((p >= 0) & (p < 64)).all(1)
    for s in range(9):
        d = min(s, D)
        cells = _cells_populated(occ, d)
        idx = np.clip(np.floor(p * (1 << d) / 64).astype(int), 0, (1 << d) - 1)
        pop = cells[idx[:, 2], idx[:, 1], idx[:, 0]]
        want_exists = inside & (pop | (s == 0))
        got_exists = rows["q"][:, s, 0] == 1
        assert np.array_equal(got_exists, want_exists), f"voxelExists scale {s}"
        want_mask = np.zeros(n, np.uint8)
        if s < D:
            sub = _cells_populated(occ, s + 1)
            for o in range(8):
                cx, cy, cz = idx[:, 0] * 2 + (o & 1), idx[:, 1] * 2 + (o >> 1 & 1), idx[:, 2] * 2 + (o >> 2 & 1)
                want_mask |= (sub[cz, cy, cx].astype(np.uint8) << o)
            want_mask = np.where(inside & (pop | (s == 0)), want_mask, 0)
        assert np.array_equal(rows["q"][:, s, 1], want_mask), f"getChildMask scale {s}"
    assert inside.sum() > 15000 and (~inside).sum() > 1000
    # LaineKarrasOctree::getStats's text (LaineKarrasOctree.cpp:336-348), default ostream formatting
    assert stats == (f"Laine-Karras SVO Statistics:\n  Total voxels: {voxel_count}\n  Max levels: 6\n"
                     f"  Memory usage: {mem / 1024.0 / 1024.0:g} MB\n  Avg bytes/voxel: {mem // voxel_count}\n")


def test_ray_hit_conversion(vx, orc, tmp_path):
    # oracle records of voxel hits, LOD stops and misses -> the adapter's RayHit, field for field
    grid, cam = vx.gen_cave(128, 5, 0.5)
    built = vx.build_from_dense(grid, bc1=True)
    scene = orc.OracleScene.from_built(built)
    rng = np.random.default_rng(8)
    n = 6000
    rays = np.zeros(n, orc.RAY_DTYPE)
    rays["o"] = rng.uniform(-40, 168, (n, 3))
    rays["d"] = rng.normal(size=(n, 3)) * rng.choice([1e-3, 1.0, 50.0], (n, 1))
    rays["d"][::17, 1] = 0.0
    rays["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, 30, n), 0.0)
    rays["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(1, 60, n), 1e30)
    plain, _ = orc.trace(scene, rays[:3000], nthreads=8)
    lod = vx.lod_with_bias(vx.lod_from_camera(1.0, 120), 5.0)
    coarse = orc.trace_lod(scene, rays[3000:], lod.ray_orig_size, lod.ray_dir_size, nthreads=8)
    hits = np.concatenate([plain, coarse])
    assert ((hits["flags"] >> 6) & 1).sum() > 100 and (hits["flags"] & 1).sum() > 2000 and (~hits["flags"] & 1).sum() > 500
    src = tmp_path / "in.bin"
    src.write_bytes(np.uint64(n).tobytes() + rays.tobytes() + hits.tobytes())
    out = tmp_path / "out.bin"
    _run("convert", src, out)
    (got_rays, got), = se.read_groups(out.read_bytes(), orc.RAY_DTYPE)
    assert got_rays.tobytes() == rays.tobytes()
    want = se.expected_rayhits(rays, hits, 7, 128.0)
    bad = se.rec_diff(got, want)
    assert len(bad) == 0, f"{len(bad)} RayHits differ, first {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"
    assert (got["entity"] == 0).all()
    # the reference's voxel hit: tMax one voxel past tMin, scale maxLevels - 1
    vox = (got["hit"] == 1) & (got["lod"] == 0)
    assert np.array_equal(got["tMax"][vox], got["tMin"][vox] + np.float32(1))
    assert (got["scale"][vox] == 6).all()
