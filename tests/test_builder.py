"""Scene builder and generators: layout, ordering, far pointers, determinism.

The descriptor order is pinned by an independent, queue-driven restatement of
LaineKarrasOctree::rebuild's ordering (libraries/SVO/src/SVORebuild.cpp:292-598:
bin voxels into bricks, Morton-sort the bricks, build parents bottom-up, BFS
from the root with each node's children appended contiguously in octant order).
The C++ builder constructs the same order level by level; the two must agree
byte for byte.
"""
from __future__ import annotations

from collections import deque

import numpy as np
import pytest


def _spread3(v: int) -> int:
    r = 0
    for b in range(21):
        r |= ((v >> b) & 1) << (3 * b)
    return r


def _morton(x, y, z) -> int:
    return _spread3(x) | (_spread3(y) << 1) | (_spread3(z) << 2)


def reference_order_build(grid: np.ndarray):
    """Queue-based restatement of the reference's build order (no far pointers)."""
    N = grid.shape[0]
    B = N // 8
    occ = (grid != 0).reshape(B, 8, B, 8, B, 8).any(axis=(1, 3, 5))  # [bz, by, bx]
    bricks = [(bx, by, bz) for bz in range(B) for by in range(B) for bx in range(B) if occ[bz, by, bx]]
    bricks.sort(key=lambda c: _morton(c[0] * 8, c[1] * 8, c[2] * 8))  # SVORebuild.cpp:363-371
    levels = B.bit_length() - 1
    # node keys (depth above bricks, coord); depth 0 = bricks
    children = {}
    cur = {c: ("brick", i) for i, c in enumerate(bricks)}
    for depth in range(1, levels + 1):
        parents = {}
        for c in cur:
            p = (c[0] // 2, c[1] // 2, c[2] // 2)
            octant = (c[0] % 2) + 2 * (c[1] % 2) + 4 * (c[2] % 2)
            parents.setdefault(p, {})[octant] = (depth - 1, c)
        for p, ch in parents.items():
            children[(depth, p)] = ch
        cur = parents
    root = (levels, (0, 0, 0))
    desc = [None]
    desc[0] = {"key": root}
    queue = deque([(root, 0)])
    out = [[0, 0]]
    while queue:
        key, idx = queue.popleft()
        ch = children.get(key, {})
        valid = 0
        for o in ch:
            valid |= 1 << o
        leaf = valid if key[0] == 1 else 0
        first = len(out)
        w0 = (valid << 16) | (leaf << 24) | (first if ch else 0)
        out[idx] = [w0, 0]
        nonleaf = [ch[o] for o in range(8) if o in ch and key[0] > 1]
        leaves = [ch[o] for o in range(8) if o in ch and key[0] == 1]
        for ck in nonleaf + leaves:
            out.append([0, 0])
        for j, ck in enumerate(nonleaf):
            queue.append((ck, first + j))
        for j, ck in enumerate(leaves):
            brick_id = bricks.index(ck[1])
            out[first + len(nonleaf) + j] = [(0xFF << 16) | (0xFF << 24), brick_id]
    return np.array(out, np.uint32), bricks


def _random_scene(N, density, seed):
    rng = np.random.default_rng(seed)
    grid = np.where(rng.random((N, N, N)) < density, rng.integers(1, 2**32, (N, N, N), dtype=np.uint64), 0)
    grid = grid.astype(np.uint32)
    # keep bricks in the two opposite root octants so the root is the full volume
    grid[0, 0, 0] = 0xFF112233
    grid[N - 1, N - 1, N - 1] = 0xFF445566
    return grid


@pytest.mark.parametrize("N,density,seed", [(16, 0.01, 1), (32, 0.002, 2), (64, 0.0005, 3), (64, 0.02, 4)])
def test_descriptor_order_matches_reference_bfs(vx, N, density, seed):
    grid = _random_scene(N, density, seed)
    built = vx.build_from_dense(grid, bc1=False)
    want, bricks = reference_order_build(grid)
    assert built.node_count < 0x8000
    assert built.nodes.tobytes() == want.tobytes()
    assert [tuple(c) for c in built.brick_coords] == bricks


def test_brick_pool_and_morton_ranks(vx, cave):
    grid, _, built = cave
    coords = built.brick_coords
    codes = np.array([_morton(*c) for c in coords[:: max(1, len(coords) // 4000)]])
    assert (np.diff(codes) > 0).all()  # strictly increasing Morton codes = brick ids are ranks
    # every populated brick appears once, and brickData is the dense grid re-gathered (z, y, x)
    B = built.N // 8
    occ = (grid != 0).reshape(B, 8, B, 8, B, 8).any(axis=(1, 3, 5))
    assert occ.sum() == built.brick_count
    for i in np.random.default_rng(0).integers(0, built.brick_count, 200):
        bx, by, bz = coords[i]
        blk = grid[bz * 8:bz * 8 + 8, by * 8:by * 8 + 8, bx * 8:bx * 8 + 8].reshape(-1)
        assert np.array_equal(built.brick_data[i], blk)
    # leaf descriptors: last brick_count entries, valid = leaf = 0xFF, word1 = brick id
    leaves = built.nodes[built.internal_count:]
    assert (leaves[:, 0] == 0xFFFF0000).all()
    assert np.array_equal(leaves[:, 1], np.arange(built.brick_count, dtype=np.uint32))


def test_bc1_blocks(vx, orc, cave):
    grid, _, built = cave
    rng = np.random.default_rng(1)
    for i in rng.integers(0, built.brick_count, 100):
        words = built.brick_data[i]
        for blk in range(32):
            texels = words[blk * 16:(blk + 1) * 16]
            block = int(built.bc1[i, blk])
            if not texels.any():
                assert block == 0  # all-empty block stored as 0 (SVORebuild.cpp:719-722)
                continue
            cols = np.stack([(texels & 255), (texels >> 8) & 255, (texels >> 16) & 255], 1) / 255.0
            cols[texels == 0] = 0.0  # empty texels are encoded as black
            assert block == vx.bc1_encode_block(cols.astype(np.float32))


def _follow(nodes, i):
    w0, w1 = int(nodes[i, 0]), int(nodes[i, 1])
    return (w1 & 0xFFFFFF) if (w0 & 0x8000) else (w0 & 0x7FFF)


def test_far_pointers_reach_every_brick(vx):
    grid, _ = vx.gen_cave(256, 11, 0.9)
    built = vx.build_from_dense(grid, bc1=False)
    nodes = built.nodes
    far = (nodes[:, 0] & 0x8000) != 0
    assert far.any() and built.node_count > 0x7FFF
    # far descriptors point beyond the 15-bit range; near ones stay within it
    assert all(_follow(nodes, i) > 0x7FFF for i in np.nonzero(far)[0])
    assert ((nodes[~far][:, 0] & 0x7FFF) <= 0x7FFF).all()
    # walking the tree reaches every leaf descriptor exactly once
    seen = np.zeros(built.node_count, bool)
    stack = [0]
    while stack:
        i = stack.pop()
        w0 = int(nodes[i, 0])
        valid, leaf = (w0 >> 16) & 255, (w0 >> 24) & 255
        if i >= built.internal_count:
            continue
        first = _follow(nodes, i)
        for k in range(bin(valid).count("1")):
            c = first + k
            assert not seen[c]
            seen[c] = True
            if not leaf:
                stack.append(c)
    assert seen[built.internal_count:].all()


def test_config_matches_reference_formulas(vx, cave, cornell):
    for _, _, built in (cave, cornell):
        c = built.config
        N = built.N
        L = int(np.log2(N))
        # VoxelSceneCacher.cpp:592-625
        assert (c.esvo_max_scale, c.user_max_levels, c.brick_depth_levels, c.brick_size) == (22, L, 3, 8)
        assert c.min_esvo_scale == 22 - L + 1
        assert c.brick_esvo_scale == 22 - (L - 1 - (L - 3))
        assert c.bricks_per_axis == N // 8
        assert list(c.grid_min) == [0, 0, 0] and list(c.grid_max) == [N, N, N]
        # localToWorld = scale(WORLD_GRID_SIZE = 10) and its inverse (VoxelSceneCacher.cpp:605-625)
        ltw = np.frombuffer(bytes(c.local_to_world), np.float32)
        wtl = np.frombuffer(bytes(c.world_to_local), np.float32)
        assert np.array_equal(ltw, np.diag([10, 10, 10, 1]).astype(np.float32).ravel())
        assert np.array_equal(wtl, np.diag(np.float32([0.1, 0.1, 0.1, 1])).ravel())


def test_generators_are_seeded(vx):
    a, ca = vx.gen_cave(64, 5, 0.3)
    b, cb = vx.gen_cave(64, 5, 0.3)
    c, _ = vx.gen_cave(64, 6, 0.3)
    assert a.tobytes() == b.tobytes() and bytes(ca) == bytes(cb)
    assert a.tobytes() != c.tobytes()
    assert abs((a != 0).mean() - 0.3) < 1e-3


def test_cave_occupancy_is_the_median(cave):
    grid = cave[0]
    assert abs((grid != 0).mean() - 0.5) < 1e-6


def test_cornell_layout(cornell):
    grid, cam, _ = cornell
    N = 64
    r, g, b = grid & 255, (grid >> 8) & 255, (grid >> 16) & 255
    assert (r[1:-1, 1:-1, 0] >= 247).all() and (g[1:-1, 1:-1, 0] <= 8).all()        # x = 0 red
    assert (g[1:-1, 1:-1, N - 1] >= 247).all() and (r[1:-1, 1:-1, N - 1] <= 8).all()  # x = N-1 green
    grey = grid[N - 1, 1:-1, 1:-1]                                                   # back wall z = N-1
    for ch in (grey & 255, (grey >> 8) & 255, (grey >> 16) & 255):
        assert ch.min() >= 192 and ch.max() <= 208
    assert (grid[1:-1, 1:-1, 1:-1] != 0).sum() == 12 ** 3 + 12 * 24 * 12
    assert list(cam.pos) == [32, 32, 0] and list(cam.dir) == [0, 0, 1] and cam.fov_deg == 60


def test_urban_small(vx):
    grid, cam = vx.gen_urban(128, 3, 1.0)
    built = vx.build_from_dense(grid, bc1=True)
    assert (grid[:, 0, :] != 0).all()    # ground plane (grid is [z, y, x], y up)
    assert (grid[:, :, :] != 0).mean() > 0.05
    assert built.brick_count > 0
    assert cam.pos[1] > 100


def test_invalid_sizes_raise(vx):
    with pytest.raises(vx.VixenError):
        vx.build_from_dense(np.zeros((24, 24, 24), np.uint32))
    with pytest.raises(vx.VixenError):
        vx.gen_cave(100, 1, 0.5)


# ----------------------------------------------------------------------------
# Sparse builder (vx_build_from_bricks)
# ----------------------------------------------------------------------------
def _bricks_of(grid):
    """Every brick cell of a dense grid[z, y, x] (empty ones included): cells (x, y, z), voxels [n, z, y, x]."""
    N = grid.shape[0]
    B = N // 8
    vox = grid.reshape(B, 8, B, 8, B, 8).transpose(0, 2, 4, 1, 3, 5).reshape(-1, 8, 8, 8)  # [bz, by, bx]
    bz, by, bx = np.meshgrid(np.arange(B), np.arange(B), np.arange(B), indexing="ij")
    cells = np.stack([bx.ravel(), by.ravel(), bz.ravel()], 1).astype(np.uint32)
    return cells, vox


def _same_built(a, b):
    assert np.array_equal(a.nodes, b.nodes)
    assert np.array_equal(a.brick_data, b.brick_data)
    assert (a.bc1 is None) == (b.bc1 is None) and (a.bc1 is None or np.array_equal(a.bc1, b.bc1))
    assert np.array_equal(a.brick_coords, b.brick_coords)
    assert bytes(a.config) == bytes(b.config) and a.internal_count == b.internal_count


@pytest.mark.parametrize("kind,N,seed", [("cave", 64, 3), ("urban", 128, 4), ("random", 32, 5), ("empty", 16, 0)])
def test_sparse_build_equals_dense(vx, kind, N, seed):
    if kind == "cave":
        grid, _ = vx.gen_cave(N, seed, 0.4)
    elif kind == "urban":
        grid, _ = vx.gen_urban(N, seed, 0.7)
    elif kind == "random":
        grid = _random_scene(N, 0.01, seed)
    else:
        grid = np.zeros((N, N, N), np.uint32)
    cells, vox = _bricks_of(grid)
    perm = np.random.default_rng(seed).permutation(len(cells))  # any order; empty bricks are dropped
    for bc1 in (True, False):
        _same_built(vx.build_from_bricks(cells[perm], vox[perm], N, bc1=bc1), vx.build_from_dense(grid, bc1=bc1))


def test_sparse_build_rejects_bad_input(vx):
    cells = np.array([[0, 0, 0], [1, 0, 0]], np.uint32)
    vox = np.full((2, 8, 8, 8), 0xFF0000FF, np.uint32)
    vx.build_from_bricks(cells, vox, 16)
    with pytest.raises(vx.VixenError, match="share a cell"):
        vx.build_from_bricks(np.array([[1, 0, 0], [1, 0, 0]], np.uint32), vox, 16)
    with pytest.raises(vx.VixenError, match="outside"):
        vx.build_from_bricks(np.array([[0, 0, 0], [2, 0, 0]], np.uint32), vox, 16)
    with pytest.raises(vx.VixenError):
        vx.build_from_bricks(cells, vox, 24)
    # a duplicate cell whose brick is empty is dropped, not an error
    dup = vx.build_from_bricks(np.array([[0, 0, 0], [1, 0, 0], [1, 0, 0]], np.uint32),
                               np.concatenate([vox, np.zeros((1, 8, 8, 8), np.uint32)]), 16)
    _same_built(dup, vx.build_from_bricks(cells, vox, 16))


@pytest.mark.parametrize("N", [2048, 4096])
def test_sparse_build_at_maximum_size(vx, N):
    """The largest grids the traversal stack supports: 8 and 9 levels above the bricks."""
    from conftest import sparse_world
    cells, vox, _ = sparse_world(N, 21, n_scattered=40, cluster=3)
    built = vx.build_from_bricks(cells, vox, N)
    B = N // 8
    assert built.config.bricks_per_axis == B and built.config.user_max_levels == B.bit_length() - 1 + 3
    assert built.brick_count == len(cells)
    # brick ids are Morton ranks of the cells, and every descriptor is in range
    codes = [_morton(int(x), int(y), int(z)) for x, y, z in built.brick_coords]
    assert codes == sorted(codes)
    from test_abi import _validate
    assert _validate(vx, built)[0] == 0
