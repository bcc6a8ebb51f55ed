"""Committed fixture (tests/golden/golden_small.npz, written by make_golden.py):
a seeded 32^3 BC1 cave with its DXT normal pool, 6072 rays and the hit records
the oracle produced for them, and the oracle's 64x48 RGBA8 images of the camera
view with the compressed shader's shading and with entry-face shading.

CPU: the oracle, the builder, the normal encoder and the generator must
reproduce the fixture bit for bit on this machine. GPU: the HIP kernels must
reproduce the stored hit records and both images.
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
    with_normals = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                                 cfg, z["normals"])
    return {"built": built, "with_normals": with_normals, "cam": cam, "rays": z["rays"].view(orc.RAY_DTYPE),
            "hits": z["hits"].view(orc.HIT_DTYPE), "counts": z["counts"], "grid_sha": z["grid_sha256"].tobytes(),
            "rgba_nrm": z["rgba_nrm"], "rgba_face": z["rgba_face"]}


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
    assert b.encode_normals().normals.tobytes() == golden["with_normals"].normals.tobytes()


def test_oracle_reproduces_fixture(orc, golden):
    hits, counts = orc.trace(orc.OracleScene.from_built(golden["built"]), golden["rays"])
    assert hits.tobytes() == golden["hits"].tobytes()
    assert counts.tolist() == golden["counts"].tolist()


def test_oracle_images_reproduce_fixture(orc, golden):
    cam = golden["cam"].as_floats()
    assert orc.render(orc.OracleScene.from_built(golden["with_normals"]), cam, 64, 48, 1e30).tobytes() == \
        golden["rgba_nrm"].tobytes()
    assert orc.render(orc.OracleScene.from_built(golden["built"]), cam, 64, 48, 1e30).tobytes() == \
        golden["rgba_face"].tobytes()
    # the two shadings differ on some hit pixels; misses (sky) are the same
    assert (golden["rgba_nrm"] != golden["rgba_face"]).any()


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


@pytest.mark.gpu
def test_hip_images_reproduce_fixture(vx, golden, gpu_ctx):
    import torch
    for key, built in (("rgba_nrm", golden["with_normals"]), ("rgba_face", golden["built"])):
        scene = vx.Scene(gpu_ctx, built)
        hits, rgba = scene.render(golden["cam"], 64, 48, tmax=1e30, rgba=True)
        torch.cuda.synchronize()
        assert vx.hits_to_numpy(hits).tobytes() == golden["hits"][:64 * 48].tobytes()
        got = rgba.cpu().numpy().view(np.uint32)
        bad = np.nonzero(got != golden[key])[0]
        assert len(bad) == 0, f"{key}: {len(bad)} pixels differ, first {bad[:4]}"
