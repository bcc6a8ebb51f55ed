"""Committed fixture (tests/golden/golden_small.npz, written by make_golden.py):
a seeded 32^3 BC1 cave, 6072 rays and the hit records the oracle produced for them.

CPU: the oracle, the builder and the generator must reproduce the fixture bit for
bit on this machine. GPU: the HIP kernels must reproduce the stored hit records.
"""
from __future__ import annotations

import ctypes as C
import hashlib
from pathlib import Path

import numpy as np
import pytest

FIXTURE = Path(__file__).with_name("golden") / "golden_small.npz"


@pytest.fixture(scope="module")
def golden(vx, orc):
    z = np.load(FIXTURE, allow_pickle=False)
    cfg = vx.VxOctreeConfig.from_buffer_copy(z["config"].tobytes())
    cam = vx.VxCamera.from_buffer_copy(z["camera"].tobytes())
    built = vx.BuiltScene(z["nodes"], z["brick_data"], z["bc1"], z["brick_coords"],
                          int(z["nodes"].shape[0] - z["brick_data"].shape[0]), cfg)
    return {"built": built, "cam": cam, "rays": z["rays"].view(orc.RAY_DTYPE), "hits": z["hits"].view(orc.HIT_DTYPE),
            "counts": z["counts"], "grid_sha": z["grid_sha256"].tobytes()}


def test_generator_and_builder_reproduce_fixture(vx, golden):
    grid, cam = vx.gen_cave(32, 9, 0.5)
    assert hashlib.sha256(grid.tobytes()).digest() == golden["grid_sha"]
    assert bytes(cam) == bytes(golden["cam"])
    b = vx.build_from_dense(grid, bc1=True)
    g = golden["built"]
    assert b.nodes.tobytes() == g.nodes.tobytes()
    assert b.brick_data.tobytes() == g.brick_data.tobytes()
    assert b.bc1.tobytes() == g.bc1.tobytes()
    assert bytes(b.config) == bytes(g.config)


def test_oracle_reproduces_fixture(orc, golden):
    hits, counts = orc.trace(orc.OracleScene.from_built(golden["built"]), golden["rays"])
    assert hits.tobytes() == golden["hits"].tobytes()
    assert counts.tolist() == golden["counts"].tolist()


def test_oracle_camera_rays_reproduce_fixture(orc, golden):
    rays = orc.camera_rays(golden["cam"].as_floats(), 64, 48, 1e30)
    assert rays.tobytes() == golden["rays"][:64 * 48].tobytes()


@pytest.mark.gpu
def test_hip_reproduces_fixture(vx, golden, gpu_ctx):
    import torch
    scene = vx.Scene(gpu_ctx, golden["built"])
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(golden["rays"])))
    torch.cuda.synchronize()
    assert got.tobytes() == golden["hits"].tobytes()
    img, _ = scene.render(golden["cam"], 64, 48, tmax=1e30)
    torch.cuda.synchronize()
    assert vx.hits_to_numpy(img).tobytes() == golden["hits"][:64 * 48].tobytes()
