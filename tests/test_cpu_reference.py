"""The oracle's traversal (D1) against a verbatim restatement of the reference's
CPU castRay (oracle/vx_cpu_reference.c: SVOTraversal.cpp:516-691 with
SVOBrickDDA.cpp:65-225,390-533) on single-level worlds, where the reference's
CPU path is valid: the worlds of its own tests (at most 2 bricks per axis).

* On the reference's scenario scenes and rays (test_ray_casting_comprehensive.cpp,
  test_brick_traversal.cpp, test_cornell_box.cpp, as ported in
  tests/test_scenarios.py and tests/test_scenarios_bricks.py) the two agree:
  same hit or miss, same hit voxel and entry face, t within 1e-5 relative once
  D1's clamp of t to the segment start (the reference clamps it to 0) is applied.
* On random rays they disagree where the reference's CPU steps are defective.
  Every disagreement is a ray where the FP64 brute force agrees with D1, and the
  restatement with its four named steps replaced by D1's (the ORC_CPU_* bits:
  no unit-cube t_max clamp, first-octant epsilon 1e-4, brick span clipped to
  the segment, brick lookup inside the child octant) agrees with D1 on every ray.
"""
from __future__ import annotations

import numpy as np
import pytest

import test_scenarios as ts
import test_scenarios_bricks as tb
from conftest import global_voxels


def _world(vx, orc, positions_or_entities):
    ents = [p if isinstance(p[0], tuple) else (p, tb.WHITE) for p in positions_or_entities]
    return tb.Scenario(vx, orc, "oracle", ents)


def _rays(specs):
    r = np.zeros(len(specs), dtype=[("o", "<f4", 3), ("tmin", "<f4"), ("d", "<f4", 3), ("tmax", "<f4")])
    for i, (o, d, t0, t1) in enumerate(specs):
        r["o"][i], r["d"][i], r["tmin"][i], r["tmax"][i] = o, d, t0, t1
    return r


def _unit(v):
    v = np.asarray(v, np.float64)
    return tuple(v / np.linalg.norm(v))


def scenario_cases():
    """(name, voxels or (position, colour) entities, [(origin, direction, tMin, tMax)]) of the ported tests."""
    X, Y, Z = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    nX, nY, nZ = (-1, 0, 0), (0, -1, 0), (0, 0, -1)
    cases = [
        ("axis_aligned", [(5, 2, 2), (2, 5, 2), (2, 2, 5), (8, 8, 8)],
         [((-5, 2, 2), X, 0, 100), ((15, 8, 8), nX, 0, 100), ((2, -5, 2), Y, 0, 100), ((8, 15, 8), nY, 0, 100),
          ((2, 2, -5), Z, 0, 100), ((8, 8, 15), nZ, 0, 100)]),
        ("diagonal", [(2, 2, 2), (5, 5, 5), (8, 8, 8), (3, 7, 4)],
         [((-2, -2, -2), _unit((1, 1, 1)), 0, 100), ((-2, 3, 5), _unit((2, 0.5, 0)), 0, 100),
          ((5, -2, -2), _unit((0, 1, 2)), 0, 100), ((-1, -1, -1), _unit((2.5, 3.7, 2.1)), 0, 100)]),
        ("room", ts._room(),
         [((5, 5, 5), X, 0, 100), ((5, 5, 5), nY, 0, 100), ((5, 5, 5), _unit((1, 1, 1)), 0, 100),
          ((2, 5, 5), X, 0, 100)]),
        ("miss", [(5, 5, 5), (2, 2, 2), (8, 8, 8)],
         [((-5, 15, 5), X, 0, 100), ((5, -5, 5), Z, 0, 100), ((15, 15, 15), (1, 1, 1), 0, 100),
          ((-1, 3.7, 3.7), X, 0, 100), ((-5, 5, 5), X, 0, 2)]),
        ("line", [(x, 5, 5) for x in range(1, 10)], [((-2, 5, 5), X, 0, 100)]),
        ("brick_hit", [(5.0, 5.0, 5.0), (5.1, 5.1, 5.1), (5.2, 5.2, 5.2)], [((-2, 5, 5), X, 0, 100)]),
        ("brick_miss", [(2.0, 2.0, 2.0), (8.0, 8.0, 8.0)], [((-2, 8, 8), X, 0, 100)]),
        ("grazing", [(2.5, 2.5, 2.5)], [((2.49, 2.0, 0.0), _unit((0.01, 0.5, 1.0)), 0, 100)]),
        ("axis_parallel", [(x, y, z) for x in np.arange(2.0, 4.01, 0.5) for y in np.arange(2.0, 4.01, 0.5)
                           for z in np.arange(2.0, 4.01, 0.5)],
         [((0, 3, 3), X, 0, 100), ((3, 0, 3), Y, 0, 100), ((3, 3, 0), Z, 0, 100)]),
        ("brick_to_brick", [(2.0, 5.0, 5.0), (7.0, 5.0, 5.0)], [((-2, 5, 5), X, 0, 100), ((-2, 5, 5), X, 5.5, 100)]),
        ("cornell_box", tb.cornell_entities(),
         [((5, 5, 5), nY, 0, 100), ((5, 5, 5), Y, 0, 100), ((5, 5, 5), nX, 0, 100), ((5, 5, 5), X, 0, 100),
          ((5, 5, 5), Z, 0, 100), ((1, 1, 1), _unit((1, 1, 1)), 0, 100), ((-5, 5, 5), nX, 0, 100),
          ((1.5, 5, 1.5), Y, 0, 100)]),
        ("small_room", tb.small_room(),
         [((4, 4, -2), Z, 0, 100), ((4, 5, -2), Z, 0, 100), ((4, 4, 4), Y, 0, 100),
          ((2, 4, 4), _unit((1, 0.2, 0.3)), 0, 100)]),
    ]
    return cases


def _compare(orc, sc, rays, ref):
    """Indices where the reference restatement's records differ from D1's (voxel, face, t after clamping)."""
    d1, _ = orc.trace(sc.oscene, rays, nthreads=8)
    gd, gr = global_voxels(sc.built, d1), global_voxels(sc.built, ref)
    same = (gd == gr).all(1)
    hit = same & (gd[:, 0] >= 0)
    face = ((d1["flags"] >> 1) & 7) == ((ref["flags"] >> 1) & 7)
    tstart = np.maximum(rays["tmin"], 0.0).astype(np.float32)  # D1 clamps t to the segment start, the reference to 0
    t_ref = np.maximum(ref["t"], tstart)
    inside = ((rays["o"] > 0) & (rays["o"] < 16)).all(1)
    t_ref = np.where(inside | (rays["tmin"] > 0), t_ref, ref["t"])
    with np.errstate(invalid="ignore"):  # misses carry t = +inf on both sides
        rel = np.abs(d1["t"].astype(np.float64) - t_ref) / np.maximum(1.0, np.abs(d1["t"].astype(np.float64)))
    bad = ~same | (hit & (~face | (rel > 1e-5)))
    return np.nonzero(bad)[0], d1


@pytest.mark.parametrize("case", scenario_cases(), ids=lambda c: c[0])
def test_reference_cpu_equals_oracle_on_the_reference_scenarios(vx, orc, case):
    name, world, specs = case
    sc = _world(vx, orc, world)
    rays = _rays(specs).view(orc.RAY_DTYPE)
    ref, status = orc.trace_cpu_reference(sc.oscene, rays)
    assert (status >= 0).all()
    bad, _ = _compare(orc, sc, rays, ref)
    assert len(bad) == 0, f"{name}: rays {bad.tolist()} differ between the reference CPU and the oracle"


RANDOM_WORLDS = {
    "cornell_box": tb.cornell_entities,
    "small_room": tb.small_room,
    "room": ts._room,
    "line": lambda: [(float(x), 5.0, 5.0) for x in range(1, 10)],
    "dense": lambda: [(2.0 + x * 0.125, 2.0 + y * 0.125, 2.0 + z * 0.125)
                      for x in range(8) for y in range(8) for z in range(8)],
}


def _random_rays(orc, rng, n):
    r = np.zeros(n, orc.RAY_DTYPE)
    r["o"] = rng.uniform(-4, 20, (n, 3))
    r["d"] = rng.normal(size=(n, 3))
    r["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, 8, n), 0.0)
    r["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(2, 20, n), 1e30)
    return r


@pytest.mark.parametrize("world", list(RANDOM_WORLDS))
def test_every_disagreement_is_a_named_reference_step(vx, orc, world):
    sc = _world(vx, orc, RANDOM_WORLDS[world]())
    rays = _random_rays(orc, np.random.default_rng(hash(world) % 2 ** 32), 60000)
    # the reference as written: where it differs from D1, the FP64 brute force sides with D1
    ref, status = orc.trace_cpu_reference(sc.oscene, rays)
    assert (status >= 0).all()
    bad, d1 = _compare(orc, sc, rays, ref)
    vox, _, amb = orc.bruteforce(sc.oscene, sc.grid, rays)
    d1_right = (global_voxels(sc.built, d1) == vox).all(1) | amb
    assert d1_right[bad].all(), f"{world}: {int((~d1_right[bad]).sum())} disagreements where D1 is not the brute force's"
    # with the four named steps replaced by D1's, the reference agrees with D1 on every ray
    adj, status = orc.trace_cpu_reference(sc.oscene, rays, orc.CPU_ALL_ADJUST)
    bad_adj, _ = _compare(orc, sc, rays, adj)
    assert len(bad_adj) == 0, f"{world}: {len(bad_adj)} rays still differ, first {rays[bad_adj[0]]}"


def test_each_named_step_matters(vx, orc):
    # dropping the unit-cube t_max, brick-clip or lookup adjustment brings back disagreements (each is
    # load-bearing); the first-octant epsilon changes no record here: position- and time-based first-octant
    # choices agree away from exact ties, and both constants are kept only because they differ
    rays = _random_rays(orc, np.random.default_rng(77), 60000)
    counts = {}
    for world in ("cornell_box", "small_room"):
        sc = _world(vx, orc, RANDOM_WORLDS[world]())
        for bit in (orc.CPU_NO_UNIT_TMAX, orc.CPU_OCTANT_EPS, orc.CPU_CLIP_BRICK, orc.CPU_LOOKUP_INSIDE):
            out, _ = orc.trace_cpu_reference(sc.oscene, rays, orc.CPU_ALL_ADJUST & ~bit)
            counts[world, bit] = len(_compare(orc, sc, rays, out)[0])
    assert counts["small_room", orc.CPU_NO_UNIT_TMAX] > 0, counts
    assert counts["cornell_box", orc.CPU_CLIP_BRICK] > 100 and counts["cornell_box", orc.CPU_LOOKUP_INSIDE] > 1000
    assert counts["cornell_box", orc.CPU_OCTANT_EPS] == counts["small_room", orc.CPU_OCTANT_EPS] == 0


def test_reference_cpu_scope(vx, orc):
    grid, _ = vx.gen_cave(32, 1, 0.5)  # 4 bricks per axis: beyond the single-level worlds
    sc = orc.OracleScene.from_built(vx.build_from_dense(grid, bc1=False))
    with pytest.raises(ValueError):
        orc.trace_cpu_reference(sc, np.zeros(1, orc.RAY_DTYPE))
