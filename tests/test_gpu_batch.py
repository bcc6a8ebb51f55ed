"""Batched persistent launch: bit-identical to per-scene launches for every launch
shape, autotuner, and hipGraph replay (each replay checked against the eager result,
with other GPU work and synchronisations between replays)."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import leaving_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_setup(vx, orc, gpu_ctx):
    import torch
    rng = np.random.default_rng(31)
    specs = [("cave", 64, 0.3), ("urban", 128, 0.8), ("cave", 128, 0.7), ("cornell", 64, 0),
             ("cave", 32, 0.5), ("urban", 64, 0.4), ("empty", 16, 0), ("cave", 256, 0.5)]
    scenes, built, rays_all, counts = [], [], [], []
    for i, (kind, N, occ) in enumerate(specs):
        if kind == "cave":
            grid, cam = vx.gen_cave(N, 100 + i, occ)
        elif kind == "urban":
            grid, cam = vx.gen_urban(N, 100 + i, occ)
        elif kind == "cornell":
            grid, cam = vx.gen_cornell(N, 100 + i)
        else:
            grid, cam = np.zeros((N, N, N), np.uint32), None
        b = vx.build_from_dense(grid, bc1=(i % 2 == 0))
        built.append(b)
        scenes.append(vx.Scene(gpu_ctx, b))
        n = int(rng.integers(0, 30000)) if kind != "empty" else 500
        if i == 3:
            n = 0  # a scene with no rays
        r = np.zeros(n, orc.RAY_DTYPE)
        if cam is not None and n:
            k = n // 2
            w = 100
            cr = orc.camera_rays(cam.as_floats(), w, (k + w - 1) // w, 1e30)[:k]
            r[:k] = cr
        r["o"][n // 2:] = rng.uniform(-0.2 * N, 1.2 * N, (n - n // 2, 3))
        r["d"][n // 2:] = rng.normal(size=(n - n // 2, 3))
        r["tmax"][n // 2:] = 1e30
        if kind in ("cave", "urban") and n > 2000:
            r[-300:] = leaving_rays(orc.RAY_DTYPE, N, 100, 200 + i)  # empty-interval exits
        rays_all.append(r)
        counts.append(n)
    offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    rays = np.concatenate(rays_all)
    d_rays = vx.rays_to_device(rays)
    # per-scene plain launches are the reference result
    plain = torch.empty(len(rays) * 32, dtype=torch.uint8, device=d_rays.device)
    for s, sc in enumerate(scenes):
        a, b = int(offsets[s]) * 32, int(offsets[s + 1]) * 32
        if b > a:
            sc.trace(d_rays[a:b], plain[a:b])
    torch.cuda.synchronize()
    batch = vx.Batch(gpu_ctx, scenes)
    return {"batch": batch, "rays": rays, "d_rays": d_rays, "offsets": offsets, "plain": plain.cpu().numpy(),
            "built": built, "scenes": scenes}


def test_plain_matches_oracle_per_scene(vx, orc, batch_setup):
    st = batch_setup
    for s, b in enumerate(st["built"]):
        lo, hi = int(st["offsets"][s]), int(st["offsets"][s + 1])
        if hi == lo:
            continue
        want, _ = orc.trace(orc.OracleScene.from_built(b), st["rays"][lo:hi], nthreads=8)
        got = st["plain"][lo * 32:hi * 32]
        assert got.tobytes() == want.tobytes(), f"scene {s}"


@pytest.mark.parametrize("tuning", [None, {"waves_per_group": 1, "chunk": 32, "lds_depth": 0, "refill_min": 1},
                                    {"waves_per_group": 2, "chunk": 64, "lds_depth": 1, "refill_min": 16},
                                    {"waves_per_group": 4, "chunk": 1024, "lds_depth": 3, "refill_min": 64},
                                    {"waves_per_group": 8, "chunk": 256, "lds_depth": 2, "refill_min": 32},
                                    {"waves_per_group": 8, "chunk": 37, "lds_depth": 4, "refill_min": 8}])
def test_batched_bit_identical_to_plain(vx, batch_setup, tuning):
    import torch
    st = batch_setup
    hits = st["batch"].trace(st["d_rays"], st["offsets"], tuning=tuning)
    torch.cuda.synchronize()
    got = hits.cpu().numpy()
    assert got.tobytes() == st["plain"].tobytes()


def test_autotune_then_default_call(vx, batch_setup):
    import torch
    st = batch_setup
    hits = torch.empty_like(st["d_rays"])
    best, ms = st["batch"].autotune(st["d_rays"], st["offsets"], hits)
    assert ms > 0 and best.waves_per_group in (1, 2, 4, 8) and 32 <= best.chunk <= 1024
    hits.zero_()
    st["batch"].trace(st["d_rays"], st["offsets"], hits)  # uses the tuned shape
    torch.cuda.synchronize()
    assert hits.cpu().numpy().tobytes() == st["plain"].tobytes()


def test_graph_replays_match_eager(vx, batch_setup):
    import torch
    st = batch_setup
    hits = torch.zeros_like(st["d_rays"])
    g = st["batch"].graph(st["d_rays"], st["offsets"], hits, tuning={"waves_per_group": 4, "chunk": 128})
    for rep in range(4):
        hits.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert hits.cpu().numpy().tobytes() == st["plain"].tobytes(), f"replay {rep}"
        # unrelated eager work between replays
        x = torch.randn(1 << 20, device=hits.device)
        (x * 2).sum().item()
    g.close()


def test_graph_keeps_its_batch(vx, gpu_ctx, batch_setup):
    """A captured graph references its batch: the batch may be destroyed first
    and the graph still replays bit-identically; the last release frees it."""
    import torch
    st = batch_setup
    batch = vx.Batch(gpu_ctx, st["scenes"])
    hits = torch.zeros_like(st["d_rays"])
    g = batch.graph(st["d_rays"], st["offsets"], hits, tuning={"waves_per_group": 2, "chunk": 256})
    batch.close()
    for _ in range(2):
        hits.zero_()
        torch.cuda.synchronize()
        g.launch()
        torch.cuda.synchronize()
        assert hits.cpu().numpy().tobytes() == st["plain"].tobytes()
    g.close()
