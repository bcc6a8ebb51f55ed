"""The batched persistent launch on the full 64-scene workload the benchmark
times (bench.batched_workload: edges drawn from {64..512}, cave and urban
generators at 10-90% occupancy, BC1, 1e4-2e6 rays per scene, 27 M rays in all,
half pinhole and half random-direction): every record equal byte for byte to
64 per-scene vx_trace launches, for the default, the autotuned and a
hand-picked launch shape, and a seeded sample of 10 000 rays per scene (all of
a scene's rays when it has fewer) equal to the oracle's records."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SAMPLE = 10_000


@pytest.fixture(scope="module")
def workload(vx, orc, gpu_ctx):
    import torch
    import bench
    expect = []

    def on_scene(s, built, rays):
        n = rays.shape[0]
        idx = np.sort(np.random.default_rng(s).choice(n, min(n, SAMPLE), replace=False))
        host = rays.cpu().numpy().view(np.uint8).reshape(-1).view(orc.RAY_DTYPE)[idx]
        want, _ = orc.trace(orc.OracleScene.from_built(built), host, nthreads=16)
        expect.append((idx, want))

    scenes, rays, offsets, edges = bench.batched_workload(vx, torch, gpu_ctx, on_scene)
    plain = torch.empty_like(rays)
    for s, sc in enumerate(scenes):
        a, b = int(offsets[s]) * 32, int(offsets[s + 1]) * 32
        sc.trace(rays[a:b], plain[a:b])
    torch.cuda.synchronize()
    batch = vx.Batch(gpu_ctx, scenes)
    yield {"scenes": scenes, "rays": rays, "offsets": offsets, "edges": edges, "plain": plain, "batch": batch,
           "expect": expect}
    batch.close()
    for sc in scenes:
        sc.close()


def test_workload_shape(workload):
    off = workload["offsets"]
    counts = np.diff(off.astype(np.int64))
    assert len(counts) == 64 and int(off[-1]) > 20_000_000
    assert counts.min() >= 10_000 and counts.max() <= 2_000_000
    assert set(workload["edges"]) == {64, 128, 256, 512}


def test_plain_launches_match_the_oracle_on_a_sample(vx, workload):
    plain = workload["plain"]
    for s, (idx, want) in enumerate(workload["expect"]):
        lo = int(workload["offsets"][s])
        got = plain[lo * 32:(lo + int(workload["offsets"][s + 1] - workload["offsets"][s])) * 32]
        got = got.view(-1, 32)[idx].cpu().numpy().reshape(-1).view(vx.HIT_DTYPE)
        assert got.tobytes() == want.tobytes(), f"scene {s}"


@pytest.mark.parametrize("tuning", [None, {"waves_per_group": 4, "chunk": 128, "lds_depth": 2, "refill_min": 16}])
def test_batched_equals_plain_byte_for_byte(workload, tuning):
    import torch
    hits = workload["batch"].trace(workload["rays"], workload["offsets"], tuning=tuning)
    torch.cuda.synchronize()
    assert torch.equal(hits, workload["plain"])


def test_autotuned_batched_equals_plain(workload):
    import torch
    hits = torch.empty_like(workload["rays"])
    best, ms = workload["batch"].autotune(workload["rays"], workload["offsets"], hits)
    assert ms > 0
    hits.zero_()
    workload["batch"].trace(workload["rays"], workload["offsets"], hits)
    torch.cuda.synchronize()
    assert torch.equal(hits, workload["plain"])
