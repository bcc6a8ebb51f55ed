"""Screen-space LOD: castRayWithLOD / castRayScreenSpaceLOD (CPU).

The LODParameters helpers behind the C ABI follow libraries/SVO/include/
SVOLOD.h:83-175 and the reference's own unit tests for them
(libraries/SVO/tests/test_lod.cpp:57-153). The oracle's LOD termination
(oracle/vx_oracle.c orc_trace_lod, after SVOTraversal.cpp:614-650) is pinned by
  - disabled LOD (dir <= 0, NaN) reproducing the plain trace record for record
    (test_lod.cpp:291-316);
  - closed forms on hand-built grids: an axis ray entering a solid grid stops at
    the coarsest child whose edge fits in the cone at the entry distance, after
    one iteration per level; a ray that first meets a valid child half way
    across the grid reports tstart + 0.5 |grid_max - grid_min| (the reference
    scales unit-cube t by the grid's diagonal length);
  - monotonicity: a wider cone never takes more iterations, and a record that
    did not stop early is the plain trace's record.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

FOV60, FOV90 = math.radians(60.0), math.radians(90.0)


def test_default_lod_is_disabled(vx):
    p = vx.VxLod()
    assert (p.ray_orig_size, p.ray_dir_size) == (0.0, 0.0) and not p.enabled()


def test_from_camera(vx):
    p = vx.lod_from_camera(FOV60, 720)
    assert p.ray_orig_size == 0.0 and p.enabled()
    pixel = float(np.float32(np.float32(FOV60) / np.float32(720.0)))
    assert p.ray_dir_size == pytest.approx(2.0 * math.tan(pixel * 0.5), rel=1e-6)
    assert vx.lod_from_camera(FOV60, 1080).ray_dir_size < p.ray_dir_size   # more pixels, narrower cone
    assert vx.lod_from_camera(FOV90, 720).ray_dir_size > p.ray_dir_size    # wider field, wider cone


def test_projected_size_and_termination(vx):
    p = vx.lod_from_camera(FOV60, 720)
    size = lambda d: d * p.ray_dir_size + p.ray_orig_size  # getProjectedPixelSize
    assert size(10.0) / size(1.0) == pytest.approx(10.0, abs=0.01)
    assert size(1000.0) >= 0.001    # far away a 1 mm voxel is below a pixel
    assert not size(1.0) >= 1.0     # nearby a 1 m voxel is not


def test_bias_and_near_plane(vx):
    base = vx.lod_from_camera(FOV60, 720, near_plane=0.25)
    pin = vx.lod_from_camera(FOV60, 720)
    assert base.ray_orig_size == pytest.approx(0.25 * pin.ray_dir_size, rel=1e-7) and base.ray_orig_size > 0
    assert base.ray_dir_size == pin.ray_dir_size
    for bias, m in ((1.0, 2.0), (-1.0, 0.5), (3.0, 8.0), (0.5, math.sqrt(2.0))):
        b = vx.lod_with_bias(base, bias)
        assert b.ray_dir_size == pytest.approx(base.ray_dir_size * m, rel=1e-6)
        assert b.ray_orig_size == pytest.approx(base.ray_orig_size * m, rel=1e-6)


def _axis_rays(orc, ys, zs, D, tmax=1e30):
    rays = np.zeros(len(ys), orc.RAY_DTYPE)
    rays["o"] = np.stack([np.full(len(ys), -D), ys, zs], 1)
    rays["d"] = [1.0, 0.0, 0.0]
    rays["tmax"] = tmax
    return rays


@pytest.fixture(scope="module")
def solid64(vx, orc):
    grid = np.full((64, 64, 64), 0xFF808080, np.uint32)
    built = vx.build_from_dense(grid, bc1=False)
    return built, orc.OracleScene.from_built(built)


@pytest.mark.parametrize("cone", [40.0, 32.0, 20.0, 16.0, 9.0, 8.0])
def test_entry_stop_level_closed_form(orc, solid64, cone):
    built, osc = solid64
    L = int(built.config.user_max_levels)
    D = 50.0
    rays = _axis_rays(orc, np.array([20.5, 3.25, 60.75]), np.array([30.5, 33.5, 0.5]), D)
    hits = orc.trace_lod(osc, rays, 0.0, cone / D)
    # children met at the entry face have t_min = 0: distance D, cone D * dir;
    # the first child of edge 2^(scale-23) * 64 <= cone stops, one iteration per level
    j = next(j for j in range(3) if 2.0 ** (-1 - j) * 64 <= cone * (1 + 1e-7))
    scale = 22 - j
    assert (hits["flags"] & 1).all() and (hits["flags"] & (1 << 6)).all()
    assert ((hits["flags"] >> 1) & 7 == 0b101).all()                 # normal (0,1,0): axis y, positive
    assert (((hits["flags"] >> 8) & 31) == scale - (23 - L)).all()   # esvoToUserScale
    assert ((hits["flags"] >> 16) == j + 1).all()
    assert np.array_equal(hits["t"], np.full(3, D, np.float32))
    assert (hits["brick"] == 0xFFFFFFFF).all() and (hits["voxel"] == 0xFFFFFFFF).all()
    assert (hits["leaf_node"][0] == 0) == (j == 0)                   # the root is the parent at scale 22
    assert (hits["leaf_node"] < built.internal_count).all()


def test_cone_below_brick_size_is_full_detail(orc, solid64):
    _, osc = solid64
    rays = _axis_rays(orc, np.array([20.5, 3.25]), np.array([30.5, 33.5]), 50.0)
    lod = orc.trace_lod(osc, rays, 0.0, 7.9 / 50.0)
    full, _ = orc.trace(osc, rays)
    assert lod.tobytes() == full.tobytes()


def test_distance_uses_grid_diagonal(vx, orc):
    grid = np.zeros((64, 64, 64), np.uint32)
    grid[:, :, 32:] = 0xFF4080C0      # solid where x >= 32
    built = vx.build_from_dense(grid, bc1=False)
    osc = orc.OracleScene.from_built(built)
    D = 10.0
    rays = _axis_rays(orc, np.array([20.5, 40.25]), np.array([30.5, 3.5]), D)
    hits = orc.trace_lod(osc, rays, 0.0, 100.0)
    # the first valid child starts half way across: unit-cube t_min = 0.5, and the
    # reference converts it with |grid_max - grid_min| = 64 sqrt(3), not the edge 64
    want = np.float32(D) + np.float32(0.5) * np.float32(math.sqrt(3 * 64.0 ** 2))
    assert (hits["flags"] & (1 << 6)).all()
    assert np.allclose(hits["t"], want, rtol=1e-6, atol=0)
    assert ((hits["flags"] >> 16) == 2).all()   # the empty first child, then the stop
    full, _ = orc.trace(osc, rays)
    assert np.allclose(full["t"], D + 32.0, rtol=1e-6)


def test_disabled_lod_is_the_plain_trace(vx, orc, cornell_bc1):
    _, cam, built = cornell_bc1
    osc = orc.OracleScene.from_built(built)
    rays = orc.camera_rays(cam.as_floats(), 80, 60, 1e30)
    full, _ = orc.trace(osc, rays)
    for dir_size in (0.0, -1.0, float("nan")):
        assert orc.trace_lod(osc, rays, 5.0, dir_size).tobytes() == full.tobytes()


def test_wider_cones_stop_sooner(vx, orc):
    grid, cam = vx.gen_cave(64, 3, 0.5)
    built = vx.build_from_dense(grid, bc1=True)
    osc = orc.OracleScene.from_built(built)
    rays = orc.camera_rays(cam.as_floats(), 96, 54, 1e30)
    full, _ = orc.trace(osc, rays)
    base = vx.lod_from_camera(math.radians(70.0), 54)
    prev = full
    fired = []
    for bias in (0.0, 2.0, 4.0, 6.0):
        p = vx.lod_with_bias(base, bias)
        lod = orc.trace_lod(osc, rays, p.ray_orig_size, p.ray_dir_size, nthreads=4)
        is_lod = (lod["flags"] & (1 << 6)) != 0
        fired.append(int(is_lod.sum()))
        assert ((lod["flags"] >> 16) <= (prev["flags"] >> 16)).all()
        assert lod[~is_lod].tobytes() == full[~is_lod].tobytes()
        # every stop satisfies the termination test at its own level
        L = int(built.config.user_max_levels)
        lvl = (lod["flags"][is_lod] >> 8) & 31
        edge = np.ldexp(1.0, (lvl.astype(np.int64) - L)) * built.N
        proj = lod["t"][is_lod].astype(np.float64) * p.ray_dir_size + p.ray_orig_size
        assert (proj >= edge * (1 - 1e-6)).all()
        prev = lod
    assert fired[0] <= fired[1] <= fired[2] <= fired[3] and fired[3] > 0


# The reference's LOD ray-casting scenarios (libraries/SVO/tests/test_lod.cpp:204-316),
# one voxel in a cubic world of the reference's size and depth, on the oracle and,
# as gpu-marked twins, through the HIP adapter's castRay*LOD methods.
BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


class _LodCaster:
    def __init__(self, vx, orc, backend, world, depth, voxel_pos):
        N = 1 << depth
        grid = np.zeros((N, N, N), np.uint32)
        x, y, z = (int(c * N / world) for c in voxel_pos)  # world position -> voxel, as the reference bins it
        grid[z, y, x] = 0xFFFFFFFF
        self.built = vx.build_from_dense(grid, bc1=False)
        for i in range(3):
            self.built.config.grid_max[i] = world  # world [0, world]^3
        self.vx, self.orc, self.backend = vx, orc, backend
        if backend == "oracle":
            self.scene = orc.OracleScene.from_built(self.built)
        else:
            self.ctx = vx.Context(0)
            self.octree = vx.LaineKarrasOctreeHip(vx.Scene(self.ctx, self.built))

    def cast(self, origin, direction, lod=None):
        """(hit, t, user scale of an LOD hit or None)"""
        if self.backend == "hip":
            h = (self.octree.castRay(origin, direction) if lod is None
                 else self.octree.castRayWithLOD(origin, direction, lod))
            return h.hit, h.tMin, h.scale
        r = np.zeros(1, self.orc.RAY_DTYPE)
        r["o"], r["d"], r["tmax"] = origin, direction, 1e30
        h = (self.orc.trace(self.scene, r)[0] if lod is None
             else self.orc.trace_lod(self.scene, r, lod.ray_orig_size, lod.ray_dir_size))[0]
        f = int(h["flags"])
        return bool(f & 1), float(h["t"]), ((f >> 8) & 31) if f & (1 << 6) else None


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_distant_voxel(vx, orc, backend):
    # DistantVoxelTerminatesEarly: one voxel at (50,50,50) in [0,100]^3, depth 8, seen
    # from 450 m with a 60 deg / 600-line camera. The cone at the grid (0.70 m) is
    # below the 3.1 m brick edge, so the cast runs to full detail: same hit as castRay.
    c = _LodCaster(vx, orc, backend, 100.0, 8, (50, 50, 50))
    full = c.cast((50.0, 50.0, 500.0), (0.0, 0.0, -1.0))
    lod = c.cast((50.0, 50.0, 500.0), (0.0, 0.0, -1.0), vx.lod_from_camera(math.radians(60.0), 600))
    assert full[0] and lod == full
    assert full[1] == pytest.approx(500.0 - 50.0 - 100.0 / 256, rel=1e-6)  # the voxel's near face
    # a cone a hundred times wider stops at a coarse node before the brick
    wide = c.cast((50.0, 50.0, 500.0), (0.0, 0.0, -1.0), vx.VxLod(0.0, 0.2))
    assert wide[0] and wide[2] is not None and wide[1] <= full[1]


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_nearby_voxel(vx, orc, backend):
    # NearbyVoxelReachesFullDetail: [0,10]^3, depth 6, voxel at (5,5,5), 3 m away, 1080 lines
    c = _LodCaster(vx, orc, backend, 10.0, 6, (5, 5, 5))
    full = c.cast((5.0, 5.0, 8.0), (0.0, 0.0, -1.0))
    lod = c.cast((5.0, 5.0, 8.0), (0.0, 0.0, -1.0), vx.lod_from_camera(math.radians(60.0), 1080))
    assert full[0] and lod == full


@pytest.mark.parametrize("backend", BACKENDS)
def test_reference_explicit_and_disabled(vx, orc, backend):
    # CastRayWithLODAcceptsExplicitParameters and DisabledLODMatchesRegularCast
    c = _LodCaster(vx, orc, backend, 10.0, 6, (5, 5, 5))
    full = c.cast((5.0, 5.0, 15.0), (0.0, 0.0, -1.0))
    explicit = c.cast((5.0, 5.0, 15.0), (0.0, 0.0, -1.0), vx.VxLod(0.0, 0.01))
    assert explicit[0]  # the cone (0.05 at the grid) is below the brick edge (1.25): full detail
    assert explicit == full
    assert c.cast((5.0, 5.0, 15.0), (0.0, 0.0, -1.0), vx.VxLod(0.0, 0.0)) == full
