"""The reference's CPU ray-casting scenario tests, ported with the same voxels, rays
and tolerances (libraries/SVO/tests/test_ray_casting_comprehensive.cpp).

The reference sizes each world from its voxels; here every scene sits in a
16^3 voxel-unit world [0,16]^3 with the voxel at position p occupying
[p, p+1]^3, which is the same world geometry, so the rays are the reference's
rays unchanged. Each scenario runs on the CPU oracle and, as a gpu-marked
twin, through the HIP library (LaineKarrasOctreeHip.castRay).
"""
from __future__ import annotations

import numpy as np
import pytest

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


class _Caster:
    def __init__(self, vx, orc, backend, voxels):
        grid = np.zeros((16, 16, 16), np.uint32)
        for p in voxels:
            x, y, z = (int(c) for c in p)  # the reference bins with static_cast<int> (SVORebuild.cpp:294-297)
            grid[z, y, x] = 0xFFFFFFFF
        self.built = vx.build_from_dense(grid, bc1=False)
        self.backend = backend
        self.orc = orc
        if backend == "oracle":
            self.scene = orc.OracleScene.from_built(self.built)
        else:
            import __graft_entry__ as entry
            vxm = entry.load_package()
            self.ctx = vxm.Context(0)
            self.octree = vxm.LaineKarrasOctreeHip(vxm.Scene(self.ctx, self.built))

    def castRay(self, origin, direction, tMin=0.0, tMax=1e30):
        if self.backend == "hip":
            h = self.octree.castRay(origin, direction, tMin, tMax)
            return h.hit, h.tMin, np.array(h.hitPoint)
        r = np.zeros(1, self.orc.RAY_DTYPE)
        r["o"], r["d"], r["tmin"], r["tmax"] = origin, direction, tMin, tMax
        h = self.orc.trace(self.scene, r)[0][0]
        if not h["flags"] & 1:
            return False, float(h["t"]), np.full(3, np.nan)
        d = np.asarray(direction, np.float64)
        hp = np.asarray(origin, np.float64) + d / np.linalg.norm(d) * float(h["t"])
        return True, float(h["t"]), hp


@pytest.fixture
def make(vx, orc, request):
    def _make(voxels):
        return _Caster(vx, orc, request.param, voxels)
    return _make


def _near(hp, voxels, tol=5.0):
    return any(np.linalg.norm(hp - np.asarray(v, np.float64)) < tol for v in voxels)


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_axis_aligned_rays_from_outside(make):
    v = [(5, 2, 2), (2, 5, 2), (2, 2, 5), (8, 8, 8)]
    o = make(v)
    for origin, d, want in (((-5, 2, 2), (1, 0, 0), v[0]), ((15, 8, 8), (-1, 0, 0), v[3]),
                            ((2, -5, 2), (0, 1, 0), v[1]), ((8, 15, 8), (0, -1, 0), v[3]),
                            ((2, 2, -5), (0, 0, 1), v[2]), ((8, 8, 15), (0, 0, -1), v[3])):
        hit, t, hp = o.castRay(origin, d, 0.0, 100.0)
        assert hit and _near(hp, [want])


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_diagonal_rays_various_angles(make):
    v = [(2, 2, 2), (5, 5, 5), (8, 8, 8), (3, 7, 4)]
    o = make(v)
    hit, t, hp = o.castRay((-2, -2, -2), _unit((1, 1, 1)), 0.0, 100.0)
    assert hit and _near(hp, v[:3])
    for origin, d in (((-2, 3, 5), (2, 0.5, 0)), ((5, -2, -2), (0, 1, 2)), ((-1, -1, -1), (2.5, 3.7, 2.1))):
        hit, t, hp = o.castRay(origin, _unit(d), 0.0, 100.0)
        if hit:
            assert _near(hp, v)


def _room():
    v = []
    for y in range(1, 10):
        for z in range(1, 10):
            v += [(1, y, z), (9, y, z)]
    for x in range(2, 9):
        for z in range(1, 10):
            v += [(x, 1, z), (x, 9, z)]
    for x in range(2, 9):
        for y in range(2, 9):
            v += [(x, y, 1), (x, y, 9)]
    return v


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_rays_from_inside_grid(make):
    o = make(_room())
    hit, t, hp = o.castRay((5, 5, 5), (1, 0, 0), 0.0, 100.0)
    assert hit and abs(hp[0] - 9.0) <= 2.0
    hit, t, hp = o.castRay((5, 5, 5), (0, -1, 0), 0.0, 100.0)
    assert hit and abs(hp[1] - 1.0) <= 2.0
    hit, t, hp = o.castRay((5, 5, 5), _unit((1, 1, 1)), 0.0, 100.0)
    assert hit and (hp > 7.0).all()
    hit, t, hp = o.castRay((2, 5, 5), (1, 0, 0), 0.0, 100.0)
    assert hit and abs(hp[0] - 9.0) <= 2.0


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_complete_miss_cases(make):
    v = [(5, 5, 5), (2, 2, 2), (8, 8, 8)]
    o = make(v)
    assert not o.castRay((-5, 15, 5), (1, 0, 0), 0.0, 100.0)[0]
    assert not o.castRay((5, -5, 5), (0, 0, 1), 0.0, 100.0)[0]
    assert not o.castRay((15, 15, 15), (1, 1, 1), 0.0, 100.0)[0]
    hit, t, hp = o.castRay((-1, 3.7, 3.7), (1, 0, 0), 0.0, 100.0)
    assert (not hit) or _near(hp, v)
    assert not o.castRay((-5, 5, 5), (1, 0, 0), 0.0, 2.0)[0]


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_multiple_voxel_traversal(make):
    o = make([(x, 5, 5) for x in range(1, 10)])
    hit, t, hp = o.castRay((-2, 5, 5), (1, 0, 0), 0.0, 100.0)
    assert hit and hp[0] < 3.0
    grid_v = [(x, y, 5) for x in range(2, 9, 2) for y in range(2, 9, 2)]
    o = make(grid_v)
    hit, t, hp = o.castRay((0, 0, 5), _unit((1, 1, 0)), 0.0, 100.0)
    assert hit and _near(hp, grid_v, 3.0)


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_dense_volume_traversal(make):
    v = [(x, y, z) for x in range(3, 8) for y in range(3, 8) for z in range(3, 8)]
    o = make(v)
    hit, t, hp = o.castRay((0, 5, 5), (1, 0, 0), 0.0, 100.0)
    assert hit and abs(hp[0] - 3.0) <= 1.0
    hit, t, hp = o.castRay((5, 5, 5), (0, 1, 0), 0.0, 100.0)  # starts inside the solid block
    assert hit and t == 0.0


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_edge_cases_and_boundaries(make):
    v = [(1, 1, 1), (9, 9, 9), (5, 5, 5)]
    o = make(v)
    hit, t, hp = o.castRay((-2, 5, 5), (1, 0, 0), 0.0, 100.0)
    assert hit and _near(hp, [v[2]])
    hit, t, hp = o.castRay((0, 0, 0), _unit((1, 1, 1)), 0.0, 100.0)
    assert hit and _near(hp, v)
    assert not o.castRay((5, 5, 5), (0, 0, 0), 0.0, 100.0)[0]
    hit, t, hp = o.castRay((-1, 5, 5), _unit((1, 0.0001, 0.0001)), 0.0, 100.0)
    assert (not hit) or _near(hp, v)


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_random_stress(make):
    """Same distributions as the reference's mt19937(42) stress test, numpy's stream."""
    rng = np.random.default_rng(42)
    v = [tuple(rng.integers(1, 10, 3)) for _ in range(50)]
    o = make(v)
    hits = misses = 0
    for _ in range(100):
        origin = rng.uniform(-5, 15, 3)
        d = rng.uniform(-1, 1, 3)
        if np.linalg.norm(d) < 0.001:
            continue
        hit, t, hp = o.castRay(origin, _unit(d), 0.0, 100.0)
        if hit:
            hits += 1
            assert 0.0 <= t <= 100.0
        else:
            misses += 1
    assert hits > 0 and misses > 0


def test_structure_queries(vx, cave):
    """getWorldMin/Max, getMaxLevels, getVoxelSize and getGPUBuffers of the castRay adapter
    (LaineKarrasOctree.h:97-108, LaineKarrasOctree.cpp:327-334), on a host-built scene."""
    grid, cam, built = cave

    class _HostScene:  # the adapter reads only the config and the host buffers
        def __init__(self, b):
            self.built, self._cfg = b, b.config

    octree = vx.LaineKarrasOctreeHip(_HostScene(built))
    assert octree.getWorldMin() == (0.0, 0.0, 0.0) and octree.getWorldMax() == (256.0, 256.0, 256.0)
    assert octree.getMaxLevels() == 8  # 5 levels above 8^3 bricks
    assert [octree.getVoxelSize(s) for s in (0, 1, 7, 8, 9)] == [256.0, 128.0, 2.0, 0.0, 0.0]
    bufs = octree.getGPUBuffers()
    assert len(bufs["hierarchyBuffer"]) == 8 * built.node_count
    assert len(bufs["attributeBuffer"]) == 2048 * built.brick_count
    assert len(bufs["compressedColorBuffer"]) == 256 * built.brick_count  # test_block_compressor.cpp:248-260
    assert bufs["compressedNormalBuffer"] == b""
