"""VIXEN's material-id colours on the host: the uncompressed shader's
getMaterialColor (shaders/Materials.glsl:18-62), the rebuild's CPU table
MaterialIdToColor (libraries/SVO/src/SVORebuild.cpp:68-120), the palette a
scene carries, and the builder's material-id mode.

Pins: the constants typed from the two sources, the mix of ids 33-39 and the
HSV wheel against float64 closed forms (within 2 float32 ulps), and the library
equal to the oracle bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

F32 = np.float32

# (id, r, g, b) as Materials.glsl and SVORebuild.cpp write them
SHARED = [(1, 1.0, 0.0, 0.0), (2, 0.0, 1.0, 0.0), (3, 0.9, 0.9, 0.9), (4, 0.85, 0.85, 0.85), (5, 0.95, 0.95, 0.95),
          (6, 0.8, 0.8, 0.8), (7, 0.4, 0.4, 0.4), (10, 0.8, 0.6, 0.2), (11, 0.6, 0.8, 0.9), (20, 1.0, 0.98, 0.9),
          (30, 0.5, 0.45, 0.4), (31, 0.6, 0.55, 0.5), (32, 0.55, 0.5, 0.45), (40, 0.8, 0.6, 0.2),
          (50, 0.2, 0.2, 0.22), (60, 0.6, 0.58, 0.55), (61, 0.7, 0.85, 0.95)]
CPU_STONE = [(33, 0.52, 0.48, 0.42), (34, 0.48, 0.44, 0.38), (35, 0.45, 0.42, 0.36), (36, 0.58, 0.52, 0.46),
             (37, 0.62, 0.56, 0.48), (38, 0.65, 0.58, 0.5), (39, 0.68, 0.6, 0.52)]
TABLE_IDS = {i for i, *_ in SHARED} | set(range(33, 40))


def _hsv64(mid: int) -> np.ndarray:
    """The HSV wheel in float64: v = 0.8, s = 0.7, hue = (id mod 256) / 256."""
    hue = (mid % 256) / 256.0
    h = hue * 6.0
    i = int(np.floor(h))
    f = h - i
    v, p, q, t = 0.8, 0.8 * 0.3, 0.8 * (1 - 0.7 * f), 0.8 * (1 - 0.7 * (1 - f))
    return np.array([(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v), (v, p, q)][i])


def _ulps(a: np.ndarray, ref64: np.ndarray) -> float:
    a = np.asarray(a, np.float32)
    return float(np.max(np.abs(a.astype(np.float64) - ref64) / np.spacing(np.abs(a).astype(np.float32)).astype(np.float64)))


IDS = list(range(256)) + [256, 257, 300, 1000, 65535, 2 ** 31 + 5, 2 ** 32 - 1]


def test_shared_constants(vx, orc):
    for mid, r, g, b in SHARED:
        want = np.array([r, g, b], F32)
        for f in (vx.material_color, vx.material_id_to_color, orc.material_color, orc.material_id_to_color):
            assert np.array_equal(f(mid), want), (f.__name__, mid)


def test_stone_ids_interpolated_by_the_shader_listed_by_the_cpu_table(vx, orc):
    a64, b64 = np.array([0.52, 0.48, 0.42]), np.array([0.68, 0.6, 0.52])
    for mid, r, g, b in CPU_STONE:
        assert np.array_equal(vx.material_id_to_color(mid), np.array([r, g, b], F32))
        assert np.array_equal(orc.material_id_to_color(mid), np.array([r, g, b], F32))
        t = (mid - 33) / 6.0
        got = vx.material_color(mid)
        assert np.array_equal(got, orc.material_color(mid))
        assert _ulps(got, a64 * (1 - t) + b64 * t) <= 2.0, mid
    # the ends of the mix are the CPU table's 33 and 39; the ids between differ
    assert np.array_equal(vx.material_color(33), vx.material_id_to_color(33))
    assert np.array_equal(vx.material_color(39), vx.material_id_to_color(39))
    assert not np.array_equal(vx.material_color(34), vx.material_id_to_color(34))


@pytest.mark.parametrize("mid", [i for i in IDS if (i % 2 ** 32) not in TABLE_IDS])
def test_hsv_wheel(vx, orc, mid):
    got = vx.material_color(mid)
    assert np.array_equal(got, orc.material_color(mid))
    assert np.array_equal(got, vx.material_id_to_color(mid))  # both tables fall back to the same wheel
    assert np.array_equal(got, orc.material_id_to_color(mid))
    assert _ulps(got, _hsv64(mid)) <= 2.0
    # saturation 0.7 and value 0.8: largest channel 0.8, smallest 0.24
    assert got.max() == F32(0.8) and abs(float(got.min()) - 0.24) < 1e-7


def test_palette(vx, orc):
    pal = vx.material_palette()
    assert pal.shape == (256, 3) and pal.dtype == np.float32
    assert np.array_equal(pal, orc.material_palette())
    for mid in (0, 1, 35, 61, 100, 255):
        assert np.array_equal(pal[mid], vx.material_color(mid))


def _material_grid(vx, N=64, seed=1):
    """A Cornell room of material ids (walls 1, 2, 3, floor 4, ceiling 5, boxes 10 and 11) with stone and
    unlisted ids sprinkled over the boxes, so every branch of both tables appears."""
    rgba, cam = vx.gen_cornell(N, seed)
    z, y, x = np.nonzero(rgba)
    ids = np.zeros_like(rgba)
    v = np.where(x == 0, 1, np.where(x == N - 1, 2, np.where(y == 0, 4, np.where(y == N - 1, 5,
                 np.where(z == N - 1, 3, np.where(x < N // 2, 10, 11))))))
    ids[z, y, x] = v
    rng = np.random.default_rng(seed)
    box = np.nonzero(ids >= 10)
    pick = rng.random(len(box[0])) < 0.3
    extra = rng.choice(np.r_[np.arange(30, 41), np.arange(100, 256), [256 + 7, 70000]], pick.sum())
    ids[box[0][pick], box[1][pick], box[2][pick]] = extra
    return ids.astype(np.uint32), cam


def test_material_mode_builder(vx):
    ids, _ = _material_grid(vx)
    built = vx.build_from_dense(ids, bc1=True, materials=True)
    plain = vx.build_from_dense(ids, bc1=False)
    assert np.array_equal(built.nodes, plain.nodes)
    assert np.array_equal(built.brick_data, plain.brick_data)  # brickData keeps the ids (VoxelSceneCacher.cpp:551-588)
    assert built.palette is None
    colours = {int(m): vx.material_id_to_color(int(m)) for m in np.unique(built.brick_data)}
    black = np.zeros(3, F32)
    for b in range(built.brick_count):
        for blk in range(32):
            w = built.brick_data[b, blk * 16:(blk + 1) * 16]
            want = 0 if not w.any() else vx.bc1_encode_block(np.stack([colours[int(x)] if x else black for x in w]))
            assert int(built.bc1[b, blk]) == want, (b, blk)
    # no BC1: the scene carries getMaterialColor's palette instead
    pal = vx.build_from_dense(ids, bc1=False, materials=True)
    assert pal.bc1 is None and np.array_equal(pal.palette, vx.material_palette())


def test_builder_rejects_an_unknown_bc1_source(vx):
    import ctypes as C
    grid = np.zeros((16, 16, 16), np.uint32)
    grid[3, 4, 5] = 7
    out = C.POINTER(vx.VxBuiltScene)()
    assert vx.LIB.vx_build_from_dense(grid.ctypes.data, 16, 3, 1, C.byref(out)) == 1
    assert b"with_bc1" in vx.LIB.vx_last_error()


def test_oracle_palette_colour(vx, orc):
    # the oracle colours a hit palette[word & 0xFF] when the scene has a palette and no BC1
    ids, cam = _material_grid(vx)
    built = vx.build_from_dense(ids, bc1=False, materials=True)
    scene = orc.OracleScene.from_built(built)
    rays = orc.camera_rays(cam.as_floats(), 96, 96, 1e4)
    hits, _ = orc.trace(scene, rays)
    h = hits[(hits["flags"] & 1) == 1]
    assert len(h) > 8000
    words = built.brick_data[h["brick"], h["voxel"]]
    assert np.array_equal(h["rgb"], built.palette[words & 0xFF])
    assert len(np.unique(words)) > 10
    # without the palette the same words decode as RGBA8
    rgba = orc.OracleScene(built.nodes, built.brick_data, None, list(built.config.grid_min),
                           list(built.config.grid_max), built.config.user_max_levels, built.config.bricks_per_axis)
    h2, _ = orc.trace(rgba, rays)
    h2 = h2[(h2["flags"] & 1) == 1]
    assert np.array_equal(h2["rgb"][:, 0], (words & 255).astype(F32) / F32(255))
