"""Pins the CPU oracle against closed forms and an independent FP64 brute force.

Nothing here re-types the oracle's own formulas: BC1 is checked against the
RGB565 bit-field closed form evaluated in numpy, camera rays against a float64
pinhole model, ray set-up against a float64 slab test, and whole traversals
against a dense-grid Amanatides-Woo walk in FP64 that never sees the octree.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import global_voxels, row_rays

F32 = np.float32


# ----------------------------------------------------------------------------
# BC1 decode
# ----------------------------------------------------------------------------
def _bc1_closed_form(head: int, sel: int) -> np.ndarray:
    """DXT1 endpoints as exact binary fractions of the RGB565 fields (numpy float32)."""
    h = np.uint32(head)
    # ref0 from bits 0-15: (head << 16) / 2^32 etc. -- exact dyadic values
    r0 = F32((int(h) & 0xFFFF) / 65536.0)
    g0 = F32((int(h) & 0x7FF) / 2048.0)
    b0 = F32((int(h) & 0x1F) / 32.0)
    # ref1 from the whole word: u32 -> f32 rounds to nearest even, then scale by 2^-32
    r1 = F32(np.float32(int(h)) * F32(2.0 ** -32))
    g1 = F32(np.float32((int(h) << 5) & 0xFFFFFFFF) * F32(2.0 ** -32))
    b1 = F32(((int(h) << 11) & 0xFFFFFFFF) / 4294967296.0)
    a, b = F32(2.0) / F32(3.0), F32(1.0) / F32(3.0)
    ref = [np.array([r0, g0, b0], F32), np.array([r1, g1, b1], F32)]
    if sel < 2:
        return ref[sel]
    w0, w1 = (a, b) if sel == 2 else (b, a)
    return (ref[0] * w0 + ref[1] * w1).astype(F32)


def test_bc1_decode_all_selectors_closed_form(orc):
    rng = np.random.default_rng(11)
    heads = list(rng.integers(0, 2**32, 400, dtype=np.uint64)) + [0, 0xFFFFFFFF, 0x0000FFFF, 0xFFFF0000, 0xF800001F]
    for head in heads:
        for sel in range(4):
            for texel in (0, 5, 15):
                bits = sel << (2 * texel)
                block = int(head) | (bits << 32)
                got = orc.bc1_decode(block, texel)
                want = _bc1_closed_form(int(head), sel)
                assert got.tobytes() == want.tobytes(), (hex(int(head)), sel, texel, got, want)


def test_bc1_endpoints_are_rgb565_fields(orc):
    """Each endpoint channel is its 5/6-bit field over 32/64 plus lower-field bits below one step."""
    rng = np.random.default_rng(12)
    for head in rng.integers(0, 2**32, 200, dtype=np.uint64):
        head = int(head)
        c0 = orc.bc1_decode(head, 0)            # selector 0 for texel 0
        c1 = orc.bc1_decode(head | (1 << 32), 0)  # selector 1
        R0, G0, B0 = (head >> 11) & 31, (head >> 5) & 63, head & 31
        R1, G1, B1 = (head >> 27) & 31, (head >> 21) & 63, (head >> 16) & 31
        for v, f, scale in ((c0[0], R0, 32), (c0[1], G0, 64), (c0[2], B0, 32),
                            (c1[0], R1, 32), (c1[1], G1, 64), (c1[2], B1, 32)):
            assert f / scale <= v + 1e-7 and v < (f + 1) / scale + 1e-7


def test_bc1_roundtrip_within_reference_tolerance(vx, orc):
    """Encoder (product) -> decoder (oracle): the reference accepts 0.05..0.35 per channel
    (libraries/VoxelData/tests/test_block_compressor.cpp:58-141)."""
    rng = np.random.default_rng(13)
    # uniform blocks reproduce within one RGB565 step
    for _ in range(50):
        c = rng.random(3).astype(F32)
        block = vx.bc1_encode_block(np.tile(c, (16, 1)))
        for t in range(16):
            assert np.abs(orc.bc1_decode(block, t) - c).max() < 0.05
    # two-colour blocks: every texel lands on an endpoint
    for _ in range(50):
        a, b = rng.random(3).astype(F32), rng.random(3).astype(F32)
        cols = np.where((np.arange(16) % 2 == 0)[:, None], a, b).astype(F32)
        block = vx.bc1_encode_block(cols)
        dec = np.stack([orc.bc1_decode(block, t) for t in range(16)])
        assert np.abs(dec - cols).max() < 0.1
    # smooth gradients stay within the reference's widest tolerance
    for _ in range(50):
        a, b = rng.random(3), rng.random(3)
        cols = np.array([a + (b - a) * (i / 15.0) for i in range(16)], F32)
        block = vx.bc1_encode_block(cols)
        dec = np.stack([orc.bc1_decode(block, t) for t in range(16)])
        assert np.abs(dec - cols).max() < 0.35


def test_rgba8_decode(orc):
    for w in (0xFF000000, 0xFF0000FF, 0xFF00FF00, 0xFFFF0000, 0xFF808080, 0x12345678):
        got = orc.rgba8_decode(w)
        want = np.array([(w & 255) / 255.0, ((w >> 8) & 255) / 255.0, ((w >> 16) & 255) / 255.0], F32)
        assert got.tobytes() == want.tobytes()


# ----------------------------------------------------------------------------
# Ray generation and set-up
# ----------------------------------------------------------------------------
def test_camera_rays_pinhole_closed_form(orc, cave):
    cam = cave[1]
    w, h = 64, 36
    rays = orc.camera_rays(cam.as_floats(), w, h, 123.0)
    pos, dirv, up, right = (np.array(x, np.float64) for x in (cam.pos, cam.dir, cam.up, cam.right))
    th = np.tan(np.radians(cam.fov_deg * 0.5))
    for py in range(h):
        for px in range(w):
            u, v = (px + 0.5) / w, (py + 0.5) / h
            d = dirv + right * (2 * u - 1) * th * cam.aspect + up * (1 - 2 * v) * th
            r = rays[py * w + px]
            got = np.array(r["d"], np.float64)
            assert np.allclose(got / np.linalg.norm(got), d / np.linalg.norm(d), atol=2e-6)
            assert np.array_equal(np.array(r["o"], np.float64), pos.astype(F32))
            assert r["tmin"] == 0.0 and r["tmax"] == F32(123.0)


def test_setup_ray_matches_float64_slab(orc, cave):
    built = cave[2]
    scene = orc.OracleScene.from_built(built)
    N = built.N
    rng = np.random.default_rng(5)
    n_hit = 0
    for _ in range(2000):
        o = rng.uniform(-0.5 * N, 1.5 * N, 3).astype(F32)
        d = rng.normal(size=3).astype(F32)
        r = np.zeros(1, orc.RAY_DTYPE)
        r["o"], r["d"], r["tmin"], r["tmax"] = o, d, 0.0, 1e30
        res = orc.setup_ray(scene, r[0])
        dn = d.astype(np.float64) / np.linalg.norm(d.astype(np.float64))
        with np.errstate(divide="ignore"):
            t0 = (0.0 - o) / dn
            t1 = (N - o) / dn
        te = np.minimum(t0, t1).max()
        tx = np.maximum(t0, t1).min()
        misses = not (te <= tx and tx >= 0)
        if misses or abs(te - tx) < 1e-3:
            assert misses == (res is None) or abs(te - tx) < 1e-3
            continue
        assert res is not None
        n_hit += 1
        _, _, du, ts, tex = res
        assert abs(ts - max(te, 0.0)) <= 1e-4 * max(1.0, abs(te))
        assert abs(tex - tx) <= 1e-4 * max(1.0, abs(tx))
        assert np.allclose(du, dn, atol=1e-6)
    assert n_hit > 300


# ----------------------------------------------------------------------------
# Whole traversals against the FP64 brute force
# ----------------------------------------------------------------------------
def _check_against_bruteforce(orc, built, grid, rays, max_amb=0.01):
    scene = orc.OracleScene.from_built(built)
    hits, _ = orc.trace(scene, rays, nthreads=8)
    vox, t, amb = orc.bruteforce(scene, grid, rays)
    got = global_voxels(built, hits)
    capped = (hits["flags"] >> 16) >= 512
    ok = (got == vox).all(1) | amb | capped
    bad = np.nonzero(~ok)[0]
    assert len(bad) == 0, f"{len(bad)} rays disagree with the FP64 brute force; first {rays[bad[0]]}: " \
                          f"oracle {got[bad[0]]} brute force {vox[bad[0]]}"
    assert amb.mean() < max_amb
    hit = (got[:, 0] >= 0) & ~amb
    rel = np.abs(hits["t"][hit].astype(np.float64) - t[hit]) / np.maximum(1.0, np.abs(t[hit]))
    assert rel.max() <= 1e-5
    # the brick id of a hit is the Morton rank of its brick cell, the hit voxel is solid
    gv = got[hit]
    assert (grid[gv[:, 2], gv[:, 1], gv[:, 0]] != 0).all()
    return hits


def test_oracle_vs_bruteforce_cornell_camera(orc, cornell):
    grid, cam, built = cornell
    rays = orc.camera_rays(cam.as_floats(), 256, 256, 1e4)
    hits = _check_against_bruteforce(orc, built, grid, rays)
    assert (hits["flags"] & 1).all()


def test_oracle_vs_bruteforce_cave_camera(orc, cave):
    grid, cam, built = cave
    rays = orc.camera_rays(cam.as_floats(), 480, 270, 1e30)
    _check_against_bruteforce(orc, built, grid, rays)


def test_oracle_vs_bruteforce_random_rays(orc, cave, cornell):
    rng = np.random.default_rng(17)
    for grid, _, built in (cave, cornell):
        N = built.N
        n = 40000
        rays = np.zeros(n, orc.RAY_DTYPE)
        rays["o"] = rng.uniform(-0.3 * N, 1.3 * N, (n, 3))
        rays["d"] = rng.normal(size=(n, 3))
        rays["tmin"] = np.where(rng.random(n) < 0.3, rng.uniform(0, N / 2, n), 0.0)
        rays["tmax"] = np.where(rng.random(n) < 0.3, rng.uniform(0, N, n), 1e30)
        _check_against_bruteforce(orc, built, grid, rays)


def test_oracle_vs_bruteforce_far_pointer_tree(vx, orc):
    grid, cam = vx.gen_cave(256, 11, 0.9)
    built = vx.build_from_dense(grid, bc1=False)
    assert (built.nodes[:, 0] & 0x8000).any()
    rays = orc.camera_rays(cam.as_floats(), 320, 180, 1e30)
    rng = np.random.default_rng(4)
    extra = np.zeros(20000, orc.RAY_DTYPE)
    extra["o"] = rng.uniform(0, 256, (20000, 3))
    extra["d"] = rng.normal(size=(20000, 3))
    extra["tmax"] = 1e30
    _check_against_bruteforce(orc, built, grid, np.concatenate([rays, extra]))


# ----------------------------------------------------------------------------
# Known answers on hand-built scenes
# ----------------------------------------------------------------------------
def _single_voxel_scene(vx, p, N=16, color=0xFF3366CC):
    grid = np.zeros((N, N, N), np.uint32)
    grid[p[2], p[1], p[0]] = color
    return grid, vx.build_from_dense(grid, bc1=False)


@pytest.mark.parametrize("axis,sign", [(a, s) for a in range(3) for s in (1, -1)])
def test_axis_ray_known_answer(vx, orc, axis, sign):
    p = [5, 9, 3]
    grid, built = _single_voxel_scene(vx, p)
    o = [p[0] + 0.5, p[1] + 0.5, p[2] + 0.5]
    o[axis] = -4.0 if sign > 0 else 20.0
    d = [0.0, 0.0, 0.0]
    d[axis] = float(sign)
    r = np.zeros(1, orc.RAY_DTYPE)
    r["o"], r["d"], r["tmax"] = o, d, 1e30
    h, counts = orc.trace(orc.OracleScene.from_built(built), r)
    h = h[0]
    assert h["flags"] & 1
    # entry face: the voxel's near plane along the axis
    near = p[axis] if sign > 0 else p[axis] + 1
    assert h["t"] == abs(near - o[axis])
    assert (h["flags"] >> 1) & 3 == axis
    assert bool((h["flags"] >> 3) & 1) == (sign < 0)  # normal points back at the ray
    lin = (p[0] % 8) + 8 * (p[1] % 8) + 64 * (p[2] % 8)
    assert h["voxel"] == lin and h["brick"] == 0
    # one brick: root, then the brick-parent chain, then the single leaf descriptor last
    assert h["leaf_node"] == built.node_count - 1
    assert np.allclose(h["rgb"], [0xCC / 255.0, 0x66 / 255.0, 0x33 / 255.0])
    assert counts[4] == 1


def test_invalid_rays_miss(orc, cave):
    scene = orc.OracleScene.from_built(cave[2])
    r = np.zeros(5, orc.RAY_DTYPE)
    r["o"] = [[1, 1, 1], [np.nan, 1, 1], [1, 1, 1], [1, 1, 1], [-10, 50, 50]]
    r["d"] = [[0, 0, 0], [1, 0, 0], [np.inf, 0, 0], [np.nan, 1, 0], [1, 0, 0]]
    r["tmax"] = [1e30, 1e30, 1e30, 1e30, 5.0]
    h, _ = orc.trace(scene, r)
    assert not (h["flags"] & 1).any()
    assert (h["flags"] >> 16 == 0).all()
    assert np.isinf(h["t"]).all() and (h["brick"] == 0xFFFFFFFF).all()


def test_empty_scene_misses(vx, orc):
    grid = np.zeros((16, 16, 16), np.uint32)
    built = vx.build_from_dense(grid)
    assert built.brick_count == 0 and built.node_count == 1
    r = np.zeros(3, orc.RAY_DTYPE)
    r["o"] = [[-1, 8, 8], [8, 8, 8], [3, 3, 3]]
    r["d"] = [[1, 0, 0], [0.3, -1, 0.2], [1, 1, 1]]
    r["tmax"] = 1e30
    h, _ = orc.trace(orc.OracleScene.from_built(built), r)
    assert not (h["flags"] & 1).any()


def test_tmin_skips_near_voxels(vx, orc):
    grid = np.zeros((16, 16, 16), np.uint32)
    grid[4, 4, 2] = 0xFF0000FF
    grid[4, 4, 10] = 0xFF00FF00
    built = vx.build_from_dense(grid, bc1=False)
    r = np.zeros(5, orc.RAY_DTYPE)
    r["o"] = [[0, 4.5, 4.5]] * 5
    r["d"] = [[1, 0, 0]] * 5
    r["tmin"] = [0.0, 5.0, 0.0, 0.0, 2.5]
    r["tmax"] = [1e30, 1e30, 1.5, 2.5, 1e30]
    h, _ = orc.trace(orc.OracleScene.from_built(built), r)
    assert h["t"][0] == 2.0 and h["t"][1] == 10.0
    assert not (h["flags"][2] & 1)               # the segment ends before the voxel
    assert h["flags"][3] & 1 and h["t"][3] == 2.0  # ...or just inside it
    assert h["flags"][4] & 1 and h["t"][4] == 2.5  # starting inside a solid voxel: t clamps to tmin


# ----------------------------------------------------------------------------
# Maximum-size sparse worlds against an FP64 slab test per solid voxel
# ----------------------------------------------------------------------------
def _sparse_bruteforce(orc, scene, solid_xyz, rays, eps=2e-5):
    """First solid voxel along each ray by an FP64 slab test against every solid
    voxel (for worlds too large for a dense grid). The segment is the oracle's
    own set-up [ts, te] in voxel units; ties within eps (shared faces, edges,
    corners, a start on a voxel face, a hit at the segment's end) are ambiguous."""
    e = lambda x: eps * max(1.0, abs(x))  # noqa: E731
    V = solid_xyz.astype(np.float64)
    n = len(rays)
    vox = np.full((n, 3), -1, np.int64)
    t = np.full(n, np.inf)
    amb = np.zeros(n, bool)
    for i in range(n):
        su = orc.setup_ray(scene, rays[i])
        if su is None:
            continue
        ol, dl, _, ts, te = su
        o = ol.astype(np.float64)
        d = np.where(np.abs(dl) < 1e-8, 0.0, dl.astype(np.float64))
        lo = np.full(len(V), -np.inf)
        hi = np.full(len(V), np.inf)
        for a in range(3):
            if d[a] != 0.0:
                t0 = (V[:, a] - o[a]) / d[a]
                t1 = (V[:, a] + 1.0 - o[a]) / d[a]
                lo = np.maximum(lo, np.minimum(t0, t1))
                hi = np.minimum(hi, np.maximum(t0, t1))
            else:
                inside = (V[:, a] <= o[a]) & (o[a] < V[:, a] + 1.0)
                lo = np.where(inside, lo, np.inf)
                hi = np.where(inside, hi, -np.inf)
        entry = np.maximum(lo, ts)
        ok = (entry <= hi) & (hi > ts) & (entry <= te)
        with np.errstate(invalid="ignore"):  # hi = -inf (a parallel axis outside the voxel): not soft
            soft = (entry <= hi + eps * np.maximum(1.0, np.abs(hi))) & (hi >= ts - e(ts)) & (entry <= te + e(te))
        if not ok.any():
            amb[i] = soft.any()
            continue
        cand = np.nonzero(ok)[0]
        b = cand[np.argmin(entry[cand])]
        tb = entry[b]
        others = soft.copy()
        others[b] = False
        amb[i] = bool((entry[others] <= tb + e(tb)).any()) or abs(tb - te) <= e(te)
        vox[i] = solid_xyz[b]
        t[i] = tb
    return vox, t, amb


@pytest.mark.parametrize("N", [2048, 4096])
def test_oracle_vs_sparse_bruteforce_maximum_size(vx, orc, N):
    from conftest import aimed_rays, sparse_world
    cells, voxels, solid = sparse_world(N, 21, n_scattered=40, cluster=3)
    built = vx.build_from_bricks(cells, voxels, N, bc1=False)
    scene = orc.OracleScene.from_built(built)
    rng = np.random.default_rng(N)
    rays = aimed_rays(orc.RAY_DTYPE, rng, 1500, N, solid)
    extra = np.zeros(500, orc.RAY_DTYPE)
    extra["o"] = rng.uniform(-0.2 * N, 1.2 * N, (500, 3))
    extra["d"] = rng.normal(size=(500, 3))
    extra["tmax"] = 3.0e38
    rays = np.concatenate([rays, extra])
    hits, _ = orc.trace(scene, rays, nthreads=8)
    vox, t, amb = _sparse_bruteforce(orc, scene, solid, rays, eps=4e-6)
    got = global_voxels(built, hits)
    ok = (got == vox).all(1) | amb
    bad = np.nonzero(~ok)[0]
    assert len(bad) == 0, f"{len(bad)} rays disagree; first {rays[bad[0]]}: oracle {got[bad[0]]} brute force {vox[bad[0]]}"
    assert amb.mean() < 0.01
    hit = (vox[:, 0] >= 0) & ~amb
    assert hit.mean() > 0.5  # the aimed rays mostly hit
    rel = np.abs(hits["t"][hit].astype(np.float64) - t[hit]) / np.maximum(1.0, np.abs(t[hit]))
    assert rel.max() <= 1e-5


# ----------------------------------------------------------------------------
# The FP64 brute force itself against an exhaustive slab test per voxel
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("N,occ", [(16, 0.05), (16, 0.3), (32, 0.02)])
def test_bruteforce_pinned_by_slab_test_per_voxel(vx, orc, N, occ):
    # orc_bruteforce (Amanatides-Woo over the dense grid) and a slab test against every solid voxel (no
    # stepping at all) must name the same first voxel and entry t on every ray neither calls ambiguous
    rng = np.random.default_rng(N * 100 + int(occ * 100))
    grid = np.where(rng.random((N, N, N)) < occ, np.uint32(0xFF336699), np.uint32(0)).astype(np.uint32)
    built = vx.build_from_dense(grid, bc1=False)
    scene = orc.OracleScene.from_built(built)
    z, y, x = np.nonzero(grid)
    solid = np.stack([x, y, z], 1)
    n = 1500
    rays = np.zeros(n, orc.RAY_DTYPE)
    rays["o"] = rng.uniform(-0.3 * N, 1.3 * N, (n, 3))
    rays["d"] = rng.normal(size=(n, 3))
    rays["d"][:100] = np.eye(3)[rng.integers(0, 3, 100)] * rng.choice([-1, 1], (100, 1))  # axis-parallel
    rays["tmin"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N / 2, n), 0.0)
    rays["tmax"] = np.where(rng.random(n) < 0.2, rng.uniform(0, N, n), 1e30)
    vox_aw, t_aw, amb_aw = orc.bruteforce(scene, grid, rays)
    vox_sl, t_sl, amb_sl = _sparse_bruteforce(orc, scene, solid, rays)
    clear = ~(amb_aw | amb_sl)
    assert clear.mean() > 0.95
    assert (vox_aw[clear] == vox_sl[clear]).all()
    hit = clear & (vox_sl[:, 0] >= 0)
    assert hit.sum() > 200
    assert np.allclose(t_aw[hit], t_sl[hit], rtol=1e-12, atol=1e-12)


def test_iteration_cap_ends_a_long_traversal(orc, brick_row):
    """The reference loop stops after 512 iterations (shaders/ESVOTraversal.glsl:25) and the ray is a miss.
    Along the brick row, rays that start close enough to the last brick reach its voxel (index 7 + 8 + 64)
    at the closed-form distance 4095 - x0 with fewer iterations the closer they start; rays from further
    back end at exactly 512 iterations with a miss record."""
    N, built = brick_row
    scene = orc.OracleScene.from_built(built)
    starts = np.array([N - 200.0, N - 800.0, N - 2000.0, N - 3000.0, 100.0, -1.0], np.float32)
    hits, _ = orc.trace(scene, row_rays(orc.RAY_DTYPE, N, starts))
    iters = hits["flags"] >> 16
    for i in range(3):
        assert hits["flags"][i] & 1 and hits["brick"][i] == 511 and hits["voxel"][i] == 79
        assert hits["t"][i] == np.float32(4095.0) - starts[i]
        assert iters[i] < 512
    assert iters[0] < iters[1] < iters[2]
    for i in range(3, 6):
        assert not hits["flags"][i] & 1 and iters[i] == 512
        assert hits["brick"][i] == 0xFFFFFFFF and np.isinf(hits["t"][i])
