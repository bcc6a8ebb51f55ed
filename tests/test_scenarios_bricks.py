"""The reference's brick-level and Cornell-box scenario tests, ported with the same
voxels, rays and tolerances:
  * libraries/SVO/tests/test_brick_traversal.cpp:74-341 (brick hit, brick miss back
    to the grid, multi-brick rays, boundary grazing, axis-parallel rays in a brick,
    dense brick, DDA step consistency, brick-to-brick transition);
  * libraries/SVO/tests/test_cornell_box.cpp:152-353 (floor, ceiling, red/green
    side walls, back wall, interior diagonal, miss, hit components, all walls);
  * libraries/SVO/tests/test_ray_casting_comprehensive.cpp:592-691 (CornellBoxScene).

World. The reference creates voxel entities at float positions; the voxel world
keys them by floor(p + 1e-5) per axis (MortonCode64::fromWorldPos,
libraries/Core/src/MortonEncoding.cpp:99-108), and a later entity at the same
key replaces the earlier one in the spatial index the brick views read
(GaiaVoxelWorld.cpp:79-96). Here each scene is a 16^3 voxel-unit world [0,16]^3
with the voxel keyed k occupying [k, k+1]^3 and the last colour written at a key
kept, so the reference's rays are used unchanged.

Each scenario runs on the CPU oracle; its gpu-marked twin runs the same rays
through the HIP library (LaineKarrasOctreeHip.castRay) and also requires the HIP
hit record to equal the oracle's bit for bit for every ray. Where the reference
asserts a strict bound at a wall voxel's face ("close to 0": hitPoint < 1 for
the floor keyed 0, which spans [0, 1]; "close to 10": hitPoint > 9 for walls
keyed 9), the port allows the face plane itself (within 1e-4, the FP32 t):
the hit is the voxel's AABB entry, which lies on that plane.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]
WHITE = (1.0, 1.0, 1.0)


def _key(p):
    return tuple(int(math.floor(np.float32(c) + np.float32(1e-5))) for c in p)


def _rgba(c):
    r, g, b = (int(round(min(max(x, 0.0), 1.0) * 255)) for x in c)
    return r | g << 8 | b << 16 | 0xFF000000


class Scenario:
    """A world of voxel entities and its castRay, on the oracle or on the GPU (checked against the oracle)."""

    def __init__(self, vx, orc, backend, entities, bc1=False, N=16):
        assert backend in ("oracle", "hip"), backend
        grid = np.zeros((N, N, N), np.uint32)
        for p, c in entities:
            x, y, z = _key(p)
            grid[z, y, x] = _rgba(c)
        self.grid = grid
        self.vx, self.orc, self.backend = vx, orc, backend
        self.built = vx.build_from_dense(grid, bc1=bc1)
        self.oscene = orc.OracleScene.from_built(self.built)
        if backend == "hip":
            self.ctx = vx.Context(0)
            self.octree = vx.LaineKarrasOctreeHip(vx.Scene(self.ctx, self.built))

    def castRay(self, origin, direction, tMin=0.0, tMax=1e30):
        r = np.zeros(1, self.orc.RAY_DTYPE)
        r["o"], r["d"], r["tmin"], r["tmax"] = origin, direction, tMin, tMax
        want = self.orc.trace(self.oscene, r)[0]
        if self.backend == "hip":
            got = self.octree.castRays(r.view(self.vx.RAY_DTYPE))
            assert got.tobytes() == want.tobytes(), f"HIP record {got[0]} != oracle {want[0]} for ray {r[0]}"
            return self.octree._cast(origin, direction, tMin, tMax, None)
        h = want[0]
        hit = bool(h["flags"] & 1)
        d = np.asarray(direction, np.float64)
        d = d / np.linalg.norm(d)
        normal = [0.0, 0.0, 0.0]
        if hit:
            normal[(int(h["flags"]) >> 1) & 3] = 1.0 if (int(h["flags"]) >> 3) & 1 else -1.0
        hp = tuple(np.asarray(origin, np.float64) + d * float(h["t"])) if hit else (math.nan,) * 3
        return self.vx.RayHit(hit, float(h["t"]), hp, tuple(normal), int(h["brick"]), int(h["voxel"]),
                              int(h["leaf_node"]), tuple(float(x) for x in h["rgb"]))


@pytest.fixture
def make(vx, orc, request):
    def _make(entities, bc1=False):
        ents = [(p, WHITE) if len(np.shape(p)) == 1 and not isinstance(p[0], tuple) else p for p in entities]
        return Scenario(vx, orc, request.param, ents, bc1)
    return _make


def _unit(v):
    v = np.asarray(v, np.float64)
    return tuple(v / np.linalg.norm(v))


EPS = 1e-4


def _near(hp, want, tol):
    return all(abs(a - b) <= tol for a, b in zip(hp, want))


# ---------------------------------------------------------------------------
# test_brick_traversal.cpp
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_hit_to_leaf_transition(make):
    o = make([(5.0, 5.0, 5.0), (5.1, 5.1, 5.1), (5.2, 5.2, 5.2)])
    h = o.castRay((-2.0, 5.0, 5.0), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and _near(h.hitPoint, (5, 5, 5), 2.0)


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_miss_returns_to_grid(make):
    o = make([(2.0, 2.0, 2.0), (8.0, 8.0, 8.0)])
    h = o.castRay((-2.0, 8.0, 8.0), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and _near(h.hitPoint, (8, 8, 8), 2.0)


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_ray_through_multiple_bricks(make):
    o = make([(float(x), 5.0, 5.0) for x in range(1, 10)])
    h = o.castRay((-2.0, 5.0, 5.0), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and abs(h.hitPoint[1] - 5.0) <= 1.0 and abs(h.hitPoint[2] - 5.0) <= 1.0
    assert h.hitPoint[0] == 1.0  # the first voxel of the line, entered at its x = 1 face


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_boundary_grazing(make):
    o = make([(2.5, 2.5, 2.5)])
    h = o.castRay((2.49, 2.0, 0.0), _unit((0.01, 0.5, 1.0)), 0.0, 100.0)
    if h.hit:  # the reference only requires a graceful result
        assert abs(h.hitPoint[0] - 2.5) <= 2.0


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_axis_parallel_rays(make):
    steps = [2.0 + 0.5 * i for i in range(5)]
    o = make([(x, y, z) for x in steps for y in steps for z in steps])
    for origin, d, (a, b) in (((0.0, 3.0, 3.0), (1, 0, 0), (1, 2)), ((3.0, 0.0, 3.0), (0, 1, 0), (0, 2)),
                              ((3.0, 3.0, 0.0), (0, 0, 1), (0, 1))):
        h = o.castRay(origin, d, 0.0, 100.0)
        assert h.hit and abs(h.hitPoint[a] - 3.0) <= 1.0 and abs(h.hitPoint[b] - 3.0) <= 1.0
        axis = 3 - a - b
        assert h.hitPoint[axis] == 2.0  # the cube of voxels 2..4 is entered at its 2 face


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_dense_brick_volume(make):
    o = make([(2.0 + x * 0.125, 2.0 + y * 0.125, 2.0 + z * 0.125)
              for x in range(8) for y in range(8) for z in range(8)])
    h = o.castRay((0.0, 2.5, 2.5), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and abs(h.hitPoint[1] - 2.5) <= 0.5 and abs(h.hitPoint[2] - 2.5) <= 0.5


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_dda_step_consistency(make):
    o = make([(3.0 + x * 0.125, 3.0 + y * 0.125, 3.0 + z * 0.125)
              for x in range(8) for y in range(8) for z in range(8) if (x + y + z) % 2 == 0])
    h = o.castRay((2.0, 3.5, 3.5), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and h.hitPoint[0] == 3.0  # the first solid voxel on the line


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_brick_to_brick_transition(make):
    o = make([(2.0, 5.0, 5.0), (7.0, 5.0, 5.0)])
    h = o.castRay((-2.0, 5.0, 5.0), (1.0, 0.0, 0.0), 0.0, 100.0)
    assert h.hit and h.hitPoint[0] == 2.0
    # past the first voxel (tMin beyond it) the ray reaches the second
    h2 = o.castRay((-2.0, 5.0, 5.0), (1.0, 0.0, 0.0), 5.5, 100.0)
    assert h2.hit and h2.hitPoint[0] == 7.0


# ---------------------------------------------------------------------------
# test_cornell_box.cpp
# ---------------------------------------------------------------------------
GREY, RED, GREEN = (0.8, 0.8, 0.8), (0.8, 0.1, 0.1), (0.1, 0.8, 0.1)


def _frange(a, b, step):
    out, x = [], np.float32(a)
    while x < b:
        out.append(float(x))
        x = np.float32(x + np.float32(step))
    return out


def cornell_entities():
    """buildCornellBox (test_cornell_box.cpp:67-138): box 10, thickness 0.2, voxel step 0.5, light radius 2;
    the walls in the reference's creation order."""
    S, T, V, L = 10.0, 0.2, 0.5, 2.0
    e = []
    for x in _frange(0, S, V):                       # floor
        for z in _frange(0, S, V):
            for y in _frange(0, T, V):
                e.append(((x, y, z), GREY))
    for x in _frange(0, S, V):                       # ceiling with the light patch
        for z in _frange(0, S, V):
            for y in _frange(S - T, S, V):
                dist = math.hypot(x - S * 0.5, z - S * 0.5)
                e.append(((x, y, z), WHITE if dist < L else GREY))
    for y in _frange(0, S, V):                       # left wall, red
        for z in _frange(0, S, V):
            for x in _frange(0, T, V):
                e.append(((x, y, z), RED))
    for y in _frange(0, S, V):                       # right wall, green
        for z in _frange(0, S, V):
            for x in _frange(S - T, S, V):
                e.append(((x, y, z), GREEN))
    for x in _frange(0, S, V):                       # back wall
        for y in _frange(0, S, V):
            for z in _frange(S - T, S, V):
                e.append(((x, y, z), GREY))
    return e


CORNELL = ["rgba8", "bc1"]


@pytest.fixture
def box(make, request):
    return make(cornell_entities(), bc1=request.param == "bc1")


def _box_params():
    out = []
    for c in CORNELL:
        out.append(pytest.param("oracle", c, id=f"oracle-{c}"))
        out.append(pytest.param("hip", c, id=f"hip-{c}", marks=pytest.mark.gpu))
    return out


@pytest.mark.parametrize("make,box", _box_params(), indirect=True)
def test_cornell_floor_ceiling_back(box):
    h = box.castRay((5.0, 5.0, 5.0), (0.0, -1.0, 0.0), 0.0, 100.0)
    assert h.hit and h.hitPoint[1] <= 1.0 + EPS and h.normal == (0.0, 1.0, 0.0)   # floor normal points up
    h = box.castRay((5.0, 5.0, 5.0), (0.0, 1.0, 0.0), 0.0, 100.0)
    assert h.hit and h.hitPoint[1] >= 9.0 - EPS and h.normal == (0.0, -1.0, 0.0)
    h = box.castRay((5.0, 5.0, 5.0), (0.0, 0.0, 1.0), 0.0, 100.0)
    assert h.hit and h.hitPoint[2] >= 9.0 - EPS


@pytest.mark.parametrize("make,box", _box_params(), indirect=True)
def test_cornell_wall_colours(box):
    h = box.castRay((5.0, 5.0, 5.0), (-1.0, 0.0, 0.0), 0.0, 100.0)    # left wall, red
    assert h.hit and h.hitPoint[0] <= 1.0 + EPS
    assert h.color[0] > 0.5 and h.color[1] < 0.3
    h = box.castRay((5.0, 5.0, 5.0), (1.0, 0.0, 0.0), 0.0, 100.0)     # right wall, green
    assert h.hit and h.hitPoint[0] >= 9.0 - EPS
    assert h.color[0] < 0.3 and h.color[1] > 0.5
    h = box.castRay((5.0, 5.0, 5.0), (0.0, 1.0, 0.0), 0.0, 100.0)     # the light patch above the centre
    assert h.hit and min(h.color) > 0.9
    h = box.castRay((1.5, 5.0, 1.5), (0.0, 1.0, 0.0), 0.0, 100.0)     # grey ceiling away from the light
    assert h.hit and all(abs(c - 0.8) < 0.05 for c in h.color)


@pytest.mark.parametrize("make,box", _box_params(), indirect=True)
def test_cornell_interior_diagonal_and_miss(box):
    h = box.castRay((1.0, 1.0, 1.0), _unit((1.0, 1.0, 1.0)), 0.0, 100.0)
    assert h.hit and max(h.hitPoint) >= 9.0 - EPS  # the far corner's walls
    h = box.castRay((-5.0, 5.0, 5.0), (-1.0, 0.0, 0.0), 0.0, 100.0)   # pointing away from the box
    assert not h.hit


@pytest.mark.parametrize("make,box", _box_params(), indirect=True)
def test_cornell_all_walls(box):
    rays = [((5, 5, 5), (0, -1, 0)), ((5, 5, 5), (0, 1, 0)), ((5, 5, 5), (-1, 0, 0)), ((5, 5, 5), (1, 0, 0)),
            ((5, 5, 5), (0, 0, 1))]
    assert sum(box.castRay(o, d, 0.0, 100.0).hit for o, d in rays) == 5


def test_cornell_voxel_keys_follow_the_reference_world(vx, orc):
    # the ceiling light patch: the last of the four positions (x, x+.5) x (z, z+.5) written at a key wins
    s = Scenario(vx, orc, "oracle", cornell_entities())
    g = s.grid
    white = _rgba(WHITE)
    for x in range(10):
        for z in range(10):
            last = math.hypot(x + 0.5 - 5.0, z + 0.5 - 5.0) < 2.0
            assert (g[z, 9, x] == white) == last, (x, z)
    assert g[5, 5, 0] == _rgba(RED) and g[5, 5, 9] == _rgba(GREEN) and g[9, 5, 5] == _rgba(GREY)
    assert g[5, 0, 0] == _rgba(RED)  # the left wall was created after the floor


# ---------------------------------------------------------------------------
# test_ray_casting_comprehensive.cpp:592-691 CornellBoxScene
# ---------------------------------------------------------------------------
def small_room():
    walls = [(x, y, 7) for x in range(1, 7) for y in range(1, 7)]
    walls += [(1, y, z) for y in range(1, 7) for z in range(1, 7)]
    walls += [(6, y, z) for y in range(1, 7) for z in range(1, 7)]
    walls += [(x, 1, z) for x in range(2, 6) for z in range(1, 7)]
    walls += [(x, 6, z) for x in range(2, 6) for z in range(1, 7)]
    objects = [(x, y, z) for x in range(3, 5) for y in range(2, 4) for z in range(3, 5)]
    return [tuple(float(c) for c in p) for p in walls + objects]


@pytest.mark.parametrize("make", BACKENDS, indirect=True)
def test_cornell_box_scene(make):
    o = make(small_room())
    h = o.castRay((4, 4, -2), (0, 0, 1), 0.0, 100.0)
    assert h.hit and h.hitPoint[2] > 2.0
    h = o.castRay((4, 5, -2), (0, 0, 1), 0.0, 100.0)
    assert h.hit and abs(h.hitPoint[2] - 7.0) <= 2.0
    h = o.castRay((4, 4, 4), (0, 1, 0), 0.0, 100.0)
    assert h.hit and abs(h.hitPoint[1] - 6.0) <= 2.0
    h = o.castRay((2, 4, 4), _unit((1, 0.2, 0.3)), 0.0, 100.0)
    assert h.hit
