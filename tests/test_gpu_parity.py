"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Integer fields (flags incl. iteration count, leaf descriptor, brick, voxel) and
the FP32 bits of t and rgb must be identical: the comparison is on the raw 32
bytes of every hit record.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import global_voxels, leaving_rays, record_diff

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _assert_same(got, want, rays=None, label=""):
    bad = record_diff(got, want)
    if len(bad):
        i = bad[0]
        msg = f"{label}: {len(bad)} of {len(want)} hit records differ; first at {i}: got {got[i]} want {want[i]}"
        if rays is not None:
            msg += f" ray {rays[i]}"
        raise AssertionError(msg)


def test_camera_rays_bit_exact(vx, orc, cornell, cave, gpu_ctx):
    for cam, w, h in ((cornell[1], 512, 512), (cave[1], 1920, 1080), (cave[1], 37, 19)):
        dev = vx.camera_rays(cam, w, h, 1e4)
        _sync()
        got = dev.cpu().numpy().view(orc.RAY_DTYPE)
        want = orc.camera_rays(cam.as_floats(), w, h, 1e4)
        assert got.tobytes() == want.tobytes(), f"camera rays differ at {w}x{h}"


@pytest.mark.parametrize("scene_name", ["cornell", "cornell_bc1"])
def test_render_cornell_512_bit_exact(vx, orc, scene_name, request, gpu_ctx):
    grid, cam, built = request.getfixturevalue(scene_name)
    scene = vx.Scene(gpu_ctx, built)
    hits, _ = scene.render(cam, 512, 512, tmax=1e4)
    _sync()
    got = vx.hits_to_numpy(hits)
    rays = orc.camera_rays(cam.as_floats(), 512, 512, 1e4)
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    _assert_same(got, want, rays, scene_name)
    assert (got["flags"] & 1).mean() > 0.99  # a closed room: every primary ray hits


def test_render_cave_1080p_bit_exact(vx, orc, cave, gpu_ctx):
    grid, cam, built = cave
    scene = vx.Scene(gpu_ctx, built)
    hits, _ = scene.render(cam, 1920, 1080, tmax=1e30)
    _sync()
    got = vx.hits_to_numpy(hits)
    rays = orc.camera_rays(cam.as_floats(), 1920, 1080, 1e30)
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    _assert_same(got, want, rays, "cave 1080p")


def _random_rays(rng, n, N, inside_frac=0.5):
    r = np.zeros(n, dtype=[("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
    k = int(n * inside_frac)
    r["o"][:k] = rng.uniform(0, N, (k, 3))
    r["o"][k:] = rng.uniform(-0.5 * N, 1.5 * N, (n - k, 3))
    r["d"] = rng.normal(size=(n, 3))
    r["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N * 0.3, n), 0.0)
    r["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N, n), 1e30)
    return r


def test_trace_random_rays_bit_exact(vx, orc, cave, cornell_bc1, gpu_ctx):
    rng = np.random.default_rng(7)
    for grid, cam, built in (cave, cornell_bc1):
        scene = vx.Scene(gpu_ctx, built)
        rays = _random_rays(rng, 300_000, built.N)
        rays = rays.astype(orc.RAY_DTYPE)
        got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
        _sync()
        want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
        _assert_same(got, want, rays, f"random rays N={built.N}")


def test_trace_edge_rays_bit_exact(vx, orc, cave, gpu_ctx):
    """Invalid, axis-parallel, grazing, outside and clipped rays."""
    grid, cam, built = cave
    N = float(built.N)
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    rows = [
        ((10, 10, 10), (0, 0, 0), 0, 1e30),          # zero direction -> miss
        ((nan, 1, 1), (1, 0, 0), 0, 1e30),           # NaN origin -> miss
        ((1, 1, 1), (inf, 0, 0), 0, 1e30),           # infinite direction -> miss
        ((-5, 100.5, 100.5), (1, 0, 0), 0, 1e30),    # axis-parallel from outside
        ((100.5, -5, 100.5), (0, 1, 0), 0, 1e30),
        ((100.5, 100.5, N + 5), (0, 0, -1), 0, 1e30),
        ((0, 0, 0), (1, 1, 1), 0, 1e30),             # corner diagonal
        ((-10, 50, 50), (1, 0, 0), 0, 5.0),          # tmax before the volume -> miss
        ((-10, 50, 50), (1, 0, 0), 500, 1e30),       # tmin after the volume -> miss
        ((N, 64, 64), (0, 1, 0), 0, 1e30),           # along the max-x face
        ((0, 64, 64), (0, 0, 1), 0, 1e30),           # along the min-x face
        ((64, 64, 64), (1e-7, 1, 1e-9), 0, 1e30),    # nearly parallel
        ((128, 128, 128), (-1, 2e-6, 3e-6), 0, 1e30),
        ((-100, -100, -100), (-1, -1, -1), 0, 1e30),  # pointing away
        ((128.0, 128.0, 128.0), (1, 0, 0), 3.0, 3.5),  # short clipped segment
    ]
    rays = np.zeros(len(rows), orc.RAY_DTYPE)
    for i, (o, d, t0, t1) in enumerate(rows):
        rays[i]["o"], rays[i]["d"], rays[i]["tmin"], rays[i]["tmax"] = o, d, t0, t1
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
    _sync()
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays)
    _assert_same(got, want, rays, "edge rays")
    assert not (got["flags"][[0, 1, 2, 7, 8, 13]] & 1).any()


@pytest.mark.parametrize("scene_name", ["cave", "cornell"])
def test_rays_leaving_through_a_max_face_bit_exact(vx, orc, scene_name, request, gpu_ctx):
    """Regression: a ray finished by initTraversalState's empty-interval exit
    must get a miss record (not its own set-up values read back as a hit)."""
    _, _, built = request.getfixturevalue(scene_name)
    rays = leaving_rays(orc.RAY_DTYPE, built.N, 2000, 17)
    osc = orc.OracleScene.from_built(built)
    want, _ = orc.trace(osc, rays, nthreads=8)
    assert (want["flags"] == 0).mean() > 0.5  # most take the empty-interval exit
    print(f"leaving rays {scene_name}", flush=True)
    scene = vx.Scene(gpu_ctx, built)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
    _sync()
    _assert_same(got, want, rays, "leaving rays")


def test_counters_match_oracle(vx, orc, cave, gpu_ctx):
    grid, cam, built = cave
    scene = vx.Scene(gpu_ctx, built)
    rays = orc.camera_rays(cam.as_floats(), 640, 360, 1e30)
    hits, counters = scene.trace_counters(vx.rays_to_device(rays))
    _sync()
    want, counts = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    _assert_same(vx.hits_to_numpy(hits), want, rays, "counters run")
    assert counters.cpu().numpy()[:5].tolist() == counts.astype(np.int64).tolist()


def test_far_pointer_scene_bit_exact(vx, orc, gpu_ctx):
    """A dense cave has more than 32767 descriptors, so child blocks use far pointers."""
    grid, cam = vx.gen_cave(256, 11, 0.9)
    built = vx.build_from_dense(grid, bc1=True)
    assert built.node_count > 0x7FFF
    assert (built.nodes[:, 0] & 0x8000).any()
    scene = vx.Scene(gpu_ctx, built)
    rng = np.random.default_rng(3)
    rays = _random_rays(rng, 200_000, built.N, inside_frac=0.3).astype(orc.RAY_DTYPE)
    got = vx.hits_to_numpy(scene.trace(vx.rays_to_device(rays)))
    _sync()
    want, _ = orc.trace(orc.OracleScene.from_built(built), rays, nthreads=8)
    _assert_same(got, want, rays, "far pointers")
    # hits land in voxels that are solid in the dense grid
    g = global_voxels(built, got)
    hit = g[:, 0] >= 0
    assert hit.any()
    assert (grid[g[hit, 2], g[hit, 1], g[hit, 0]] != 0).all()


def _rgba_vs_oracle(vx, orc, gpu_ctx, built, cam, w, h, tmax, label):
    scene = vx.Scene(gpu_ctx, built)
    _, img = scene.render(cam, w, h, tmax=tmax, hits=False, rgba=True)
    _sync()
    got = img.cpu().numpy().view(np.uint32)
    want = orc.render(orc.OracleScene.from_built(built), cam.as_floats(), w, h, tmax, nthreads=16)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, f"{label}: {len(bad)} of {w * h} RGBA8 pixels differ, first {bad[:4]}"
    assert ((got >> 24) == 255).all()
    return got


@pytest.mark.parametrize("scene_name", ["cornell", "cornell_bc1"])
def test_render_rgba_entry_face_shading_bit_exact(vx, orc, scene_name, request, gpu_ctx):
    # no normal pool: computeLighting with the hit record's colour and entry-face normal
    grid, cam, built = request.getfixturevalue(scene_name)
    _rgba_vs_oracle(vx, orc, gpu_ctx, built, cam, 512, 512, 1e4, scene_name)


def test_render_rgba_compressed_shading_bit_exact(vx, orc, cornell_bc1, cave, gpu_ctx):
    # BC1 colours + DXT normals: the compressed shader's colour decode and normal blend
    grid, cam, built = cornell_bc1
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    _rgba_vs_oracle(vx, orc, gpu_ctx, b, cam, 512, 512, 1e4, "cornell+normals")
    grid, cam, built = cave
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    img = _rgba_vs_oracle(vx, orc, gpu_ctx, b, cam, 1920, 1080, 1e30, "cave+normals")
    # the normal pool changes the shading of some hit pixels
    face = _rgba_vs_oracle(vx, orc, gpu_ctx, built, cam, 1920, 1080, 1e30, "cave")
    assert (img != face).mean() > 0.05


@pytest.mark.parametrize("lds_depth", [1, 2, 4])
def test_render_with_staged_descriptors_bit_exact(vx, orc, cave, gpu_ctx, lds_depth):
    """vx_render's LDS-staged variants (the top breadth-first levels read from LDS):
    hit records, entry-face RGBA8 and compressed-shader RGBA8, all against the oracle."""
    grid, cam, built = cave
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    W, H = 640, 360
    rays = orc.camera_rays(cam.as_floats(), W, H, 1e30)
    osc = orc.OracleScene.from_built(b)
    want_hits, _ = orc.trace(osc, rays, nthreads=16)
    want_img = orc.render(osc, cam.as_floats(), W, H, 1e30, nthreads=16)
    want_face = orc.render(orc.OracleScene.from_built(built), cam.as_floats(), W, H, 1e30, nthreads=16)
    print(f"staged render lds_depth={lds_depth}", flush=True)
    for scene_built, want_rgba, label in ((b, want_img, "compressed"), (built, want_face, "entry face")):
        scene = vx.Scene(gpu_ctx, scene_built)
        scene.set_lds_depth(lds_depth)
        hits, img = scene.render(cam, W, H, tmax=1e30, hits=True, rgba=True)
        _sync()
        _assert_same(vx.hits_to_numpy(hits), want_hits, rays, f"staged {lds_depth} {label} hits")
        got = img.cpu().numpy().view(np.uint32)
        assert np.array_equal(got, want_rgba), f"staged {lds_depth} {label}: {(got != want_rgba).sum()} pixels differ"
        hits_only, _ = scene.render(cam, W, H, tmax=1e30, hits=True, rgba=False)
        _sync()
        _assert_same(vx.hits_to_numpy(hits_only), want_hits, rays, f"staged {lds_depth} hits only")
        scene.close()
    # vx_trace and vx_trace_counters read the same staged prefix
    scene = vx.Scene(gpu_ctx, built)
    scene.set_lds_depth(lds_depth)
    d_rays = vx.rays_to_device(rays)
    traced = vx.hits_to_numpy(scene.trace(d_rays))
    counted, counters = scene.trace_counters(d_rays)
    _sync()
    _assert_same(traced, want_hits, rays, f"staged {lds_depth} vx_trace")
    _assert_same(vx.hits_to_numpy(counted), want_hits, rays, f"staged {lds_depth} vx_trace_counters")
    _, counts = orc.trace(osc, rays, nthreads=16)
    assert counters.cpu().numpy()[:5].tolist() == counts.astype(np.int64).tolist()
    scene.close()


def test_scene_over_device_buffers_renders_like_copied(vx, orc, cave, gpu_ctx):
    """vx_scene_create_device: the scene buffers as CUDA tensors, used in place."""
    import torch
    grid, cam, built = cave
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    dev = torch.device("cuda", 0)
    t = {k: torch.from_numpy(getattr(b, k).view(np.uint32).copy()).to(dev) for k in ("nodes", "brick_data", "bc1",
                                                                                    "normals")}
    scene = vx.Scene.from_device(gpu_ctx, t["nodes"], t["brick_data"], b.config, bc1=t["bc1"], normals=t["normals"])
    ref = vx.Scene(gpu_ctx, b)
    h1, i1 = scene.render(cam, 640, 360, tmax=1e30, rgba=True)
    h2, i2 = ref.render(cam, 640, 360, tmax=1e30, rgba=True)
    _sync()
    assert vx.hits_to_numpy(h1).tobytes() == vx.hits_to_numpy(h2).tobytes()
    assert i1.cpu().numpy().tobytes() == i2.cpu().numpy().tobytes()
    # a batch copies the borrowed pools, so it outlives the scene
    batch = vx.Batch(gpu_ctx, [scene])
    scene.close()
    rays = vx.camera_rays(cam, 640, 360, 1e30)
    got = batch.trace(rays, np.array([0, 640 * 360], np.uint64))
    _sync()
    assert vx.hits_to_numpy(got).tobytes() == vx.hits_to_numpy(h2).tobytes()
    # misaligned brick words are refused before any device access
    with pytest.raises(vx.VixenError):
        vx.Scene.from_device(gpu_ctx, t["nodes"], t["brick_data"][1:], b.config)


def test_vixen_layout_cache_file_traces_like_the_oracle(vx, orc, cave, gpu_ctx, tmp_path):
    # A VSCA v1 file laid out as VoxelSceneCacher::SerializeToFile writes it, assembled field by field
    # (tests/vsca_layout.py: 260-byte OctreeConfig with worldGridSize = 10 at +256, scale(10) matrices,
    # gridMax = N), loads and traces bit-exact vs the oracle on the loaded buffers: in voxel units (the
    # config's grid box) and in the compute shader's 10-unit frame.
    import vsca_layout as L
    grid, cam, built = cave
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    f = tmp_path / "cave.vsca"
    f.write_bytes(L.file([L.built_entry(vx, b, 2, 256, 0.5, 2)]))
    (meta, loaded), = vx.read_cache(f)
    assert meta["world_grid_size"] == 10.0 and list(loaded.config.grid_max) == [256.0] * 3
    w, h = 640, 360
    scene = vx.Scene(gpu_ctx, loaded)
    hits, img = scene.render(cam, w, h, tmax=1e30, rgba=True)
    _sync()
    osc = orc.OracleScene.from_built(loaded)
    rays = orc.camera_rays(cam.as_floats(), w, h, 1e30)
    want, _ = orc.trace(osc, rays)
    got = vx.hits_to_numpy(hits)
    assert len(record_diff(got, want)) == 0
    assert np.array_equal(img.cpu().numpy().view(np.uint32), orc.render(osc, cam.as_floats(), w, h, 1e30))
    # the shader's frame: world = 10 / N voxel units; the camera moves with it
    k = np.float32(10.0) / np.float32(256.0)
    sf = vx.shader_frame(loaded, meta["world_grid_size"])
    cam10 = vx.VxCamera.from_buffer_copy(bytes(cam))
    cam10.pos[:] = [float(np.float32(p) * k) for p in cam.pos]
    scene10 = vx.Scene(gpu_ctx, sf)
    h10, _ = scene10.render(cam10, w, h, tmax=1e30)
    _sync()
    want10, _ = orc.trace(orc.OracleScene.from_built(sf), orc.camera_rays(cam10.as_floats(), w, h, 1e30))
    got10 = vx.hits_to_numpy(h10)
    assert len(record_diff(got10, want10)) == 0
    # the same surfaces: hit voxels agree with the voxel-frame render except where a ray grazes an edge
    same = (got10["brick"] == got["brick"]) & (got10["voxel"] == got["voxel"])
    assert same.mean() > 0.995
    assert np.allclose(got10["t"][same & (got["flags"] & 1 == 1)] / k,
                       got["t"][same & (got["flags"] & 1 == 1)], rtol=1e-4)


def test_scene_from_cache_file_renders_like_built(vx, orc, cave, gpu_ctx, tmp_path):
    # a VSCA v1 cache round trip (VIXEN's scene cache format) feeds vx_scene_create unchanged
    grid, cam, built = cave
    b = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                      built.config).encode_normals()
    f = tmp_path / "cave.vsca"
    vx.write_cache(f, [dict(built=b, scene_type="tunnels", resolution=256, density=0.5, seed=2)])
    (_, loaded), = vx.read_cache(f)
    h1, i1 = vx.Scene(gpu_ctx, b).render(cam, 640, 360, tmax=1e30, rgba=True)
    h2, i2 = vx.Scene(gpu_ctx, loaded).render(cam, 640, 360, tmax=1e30, rgba=True)
    _sync()
    assert vx.hits_to_numpy(h1).tobytes() == vx.hits_to_numpy(h2).tobytes()
    assert i1.cpu().numpy().tobytes() == i2.cpu().numpy().tobytes()
