"""VIXEN's material-id scenes on the GPU, bit-exact against the oracle.

A Cornell room of material ids (tests/test_materials.py::_material_grid) is
traced and rendered
* with getMaterialColor's palette and no BC1 pool: the uncompressed pipeline's
  colour, palette[brickData & 0xFF] (shaders/VoxelRayMarch.comp:189-191);
* with BC1 blocks encoded from MaterialIdToColor: the compressed pipeline's;
* after a round trip through a VSCA v1 cache file (written as VIXEN lays it
  out, brickData holding the ids), both ways;
* in a batch beside an RGBA8 scene (one launch, per-scene palettes).
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import record_diff
from test_materials import _material_grid

pytestmark = pytest.mark.gpu


def _check_trace(vx, orc, ctx, built, rays, label):
    scene = vx.Scene(ctx, built)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    bad = record_diff(got, want)
    assert len(bad) == 0, f"{label}: {len(bad)} of {len(rays)} records differ, first at {bad[0]}"
    return scene, got


@pytest.fixture(scope="module")
def material_scene(vx):
    ids, cam = _material_grid(vx)
    return ids, cam


def test_palette_and_bc1_from_materials(vx, orc, gpu_ctx, material_scene):
    import torch
    ids, cam = material_scene
    rays = orc.camera_rays(cam.as_floats(), 320, 320, 1e4)
    pal = vx.build_from_dense(ids, bc1=False, materials=True)
    scene, got = _check_trace(vx, orc, gpu_ctx, pal, rays, "palette")
    hit = (got["flags"] & 1) == 1
    assert hit.mean() > 0.95
    words = pal.brick_data[got["brick"][hit], got["voxel"][hit]]
    assert np.array_equal(got["rgb"][hit], vx.material_palette()[words & 0xFF])
    # the face-shaded frame of the uncompressed pipeline
    _, rgba = scene.render(cam, 320, 320, tmax=1e4, rgba=True)
    torch.cuda.synchronize()
    want = orc.render(orc.OracleScene.from_built(pal), cam.as_floats(), 320, 320, 1e4)
    assert np.array_equal(rgba.cpu().numpy().reshape(-1).view(np.uint32), want)
    comp = vx.build_from_dense(ids, bc1=True, materials=True)
    _, got_c = _check_trace(vx, orc, gpu_ctx, comp, rays, "bc1 from materials")
    # BC1 colours take precedence over a palette when a scene has both
    both = vx.BuiltScene(comp.nodes, comp.brick_data, comp.bc1, comp.brick_coords, comp.internal_count,
                         comp.config, None, vx.material_palette())
    _, got_b = _check_trace(vx, orc, gpu_ctx, both, rays, "bc1 and palette")
    assert record_diff(got_b, got_c).size == 0


def test_vsca_file_with_material_ids(vx, orc, gpu_ctx, material_scene, tmp_path):
    ids, cam = material_scene
    comp = vx.build_from_dense(ids, bc1=True, materials=True)
    path = tmp_path / "cornell_materials.cache"
    vx.write_cache(path, [{"built": comp, "scene_type": "cornell", "resolution": 64, "density": 0.1, "seed": 1}])
    (info, back), = vx.read_cache(path)
    assert np.array_equal(back.brick_data, comp.brick_data) and np.array_equal(back.bc1, comp.bc1)
    rays = orc.camera_rays(cam.as_floats(), 256, 256, 1e4)
    _check_trace(vx, orc, gpu_ctx, back, rays, "cache file, compressed colours")
    # the same file through the uncompressed pipeline: no BC1, getMaterialColor's palette
    unc = vx.BuiltScene(back.nodes, back.brick_data, None, back.brick_coords, back.internal_count, back.config,
                        None, vx.material_palette())
    _, got = _check_trace(vx, orc, gpu_ctx, unc, rays, "cache file, palette")
    hit = (got["flags"] & 1) == 1
    assert len(np.unique(got["rgb"][hit], axis=0)) > 10


def test_batch_with_palette_scenes(vx, orc, gpu_ctx, material_scene):
    ids, cam = material_scene
    pal = vx.build_from_dense(ids, bc1=False, materials=True)
    grid, cam2 = vx.gen_cornell(64, 2)
    rgba = vx.build_from_dense(grid, bc1=False)
    scenes = [vx.Scene(gpu_ctx, pal), vx.Scene(gpu_ctx, rgba), vx.Scene(gpu_ctx, pal)]
    r1 = orc.camera_rays(cam.as_floats(), 128, 128, 1e4)
    r2 = orc.camera_rays(cam2.as_floats(), 100, 90, 1e4)
    rays = np.concatenate([r1, r2, r1[::3]])
    offs = np.array([0, len(r1), len(r1) + len(r2), len(rays)], np.uint64)
    batch = vx.Batch(gpu_ctx, scenes)
    got = vx.hits_to_numpy(batch.trace(vx.rays_to_device(rays), offs))
    for s, built in enumerate((pal, rgba, pal)):
        sub = rays[int(offs[s]):int(offs[s + 1])]
        want, _ = orc.trace(orc.OracleScene.from_built(built), sub, nthreads=8)
        bad = record_diff(got[int(offs[s]):int(offs[s + 1])], want)
        assert len(bad) == 0, f"scene {s}: {len(bad)} records differ"
