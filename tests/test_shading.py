"""The compressed shader's normal pool and shading (CPU).

Normal encoder (vbvs--vixen_amd/csrc/vx_normals.cpp) and the oracle's decoders
and per-pixel shading (oracle/vx_oracle.c orc_shade) are pinned against:
  - the reference's own DXT normal tests (libraries/VoxelData/tests/
    test_block_compressor.cpp:159-206), with their tolerances;
  - encode -> decode round trips (a transposed or sign-flipped component in
    either direction fails them);
  - closed forms of the geometric normal (SVORebuild.cpp:173-193) on hand-built
    bricks;
  - the shader's BC1 decode against the CPU decode, which is pinned exhaustively
    elsewhere: they may differ only in the last bits (Compression.glsl:60-75 vs
    DXT1Compressor.cpp:14-36);
  - a float64 evaluation of the whole shading formula per pixel (colour, normal
    blend, light, sky), within one 8-bit step.
"""
from __future__ import annotations

import numpy as np
import pytest


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def _decode_block(orc, a, b):
    return np.array([orc.dxt_normal_decode(a, b, t) for t in range(16)], np.float64)


def test_reference_up_vector(vx, orc):
    # DXTNormalUpVector: all normals up -> every decoded normal has y > 0.9
    a, b = vx.dxt_normal_encode_block(np.tile([0.0, 1.0, 0.0], (16, 1)))
    out = _unit(_decode_block(orc, a, b))
    assert (out[:, 1] > 0.9).all()


def test_reference_varied(vx, orc):
    # DXTNormalVaried: coherent upward normals, >= 8 of 16 within dot > 1 - 0.5
    ang = np.arange(16, dtype=np.float32) * np.float32(0.4)
    n = _unit(np.stack([np.cos(ang) * 0.3, np.full(16, 0.9), np.sin(ang) * 0.3], 1)).astype(np.float32)
    a, b = vx.dxt_normal_encode_block(n)
    out = _unit(_decode_block(orc, a, b))
    assert ((out * n).sum(1) > 0.5).sum() >= 8


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_uniform_blocks_round_trip(vx, orc, axis, sign):
    # 16 equal normals near each axis and sign: the base normal carries them; the
    # decoded direction is within the raw format's 14-bit resolution
    rng = np.random.default_rng(axis * 2 + (sign > 0))
    for _ in range(20):
        n = np.zeros(3)
        n[axis] = sign
        n += rng.uniform(-0.6, 0.6, 3) * (np.arange(3) != axis)
        n = _unit(n).astype(np.float32)
        a, b = vx.dxt_normal_encode_block(np.tile(n, (16, 1)))
        out = _unit(_decode_block(orc, a, b))
        assert np.abs(out - n).max() < 2e-3, (n, out[0])


def test_mixed_blocks_round_trip(vx, orc):
    # two clusters of normals (a surface edge): every decoded normal is near its input
    rng = np.random.default_rng(7)
    worst = []
    for _ in range(50):
        c0, c1 = _unit(rng.normal(size=3)), _unit(rng.normal(size=3))
        if c0 @ c1 < 0.2:
            c1 = _unit(c0 + 0.8 * c1)
        pick = rng.random(16) < 0.5
        n = _unit(np.where(pick[:, None], c0, c1) + rng.normal(scale=0.02, size=(16, 3))).astype(np.float32)
        a, b = vx.dxt_normal_encode_block(n)
        out = _unit(_decode_block(orc, a, b))
        worst.append(((out * n).sum(1)).min())
    assert np.median(worst) > 0.95 and min(worst) > 0.7


def test_raw_normal_bit_layout(orc):
    # decodeRawNormal: bit 31 sign, bits 29-30 dominant axis, U = bits 14-28 and
    # V = bits 0-13 as signed fields, placed as (t,u,v) / (v,t,u) / (u,v,t)
    rng = np.random.default_rng(3)
    for axis in range(3):
        for sign in (0, 1):
            U = int(rng.integers(-0x4000, 0x4000))
            V = int(rng.integers(-0x2000, 0x2000))
            head = (sign << 31) | (axis << 29) | ((U & 0x7FFF) << 14) | (V & 0x3FFF)
            n = orc.dxt_normal_decode(head, 0, 0).astype(np.float64)  # U/V axes zero: the base itself
            t = -2.0 ** 31 if sign else 2.0 ** 31
            u = U * 2.0 ** 17 + (V & 0x3FFF) * 8.0  # int(value << 3): V's bits ride below U
            v = V * 2.0 ** 18
            tuv = {0: (t, u, v), 1: (v, t, u), 2: (u, v, t)}[axis]
            assert np.allclose(n, tuv, rtol=2 ** -23, atol=0), (axis, sign, n, tuv)


def test_geometric_normals_closed_forms(vx):
    b = np.zeros((8, 8, 8), np.uint32)  # [z, y, x]
    b[:, :4, :] = 1                      # a floor 4 voxels thick
    flat = b.reshape(512)
    assert np.allclose(vx.geometric_normal(flat, 3, 3, 3), [0, 1, 0])      # top surface
    assert np.allclose(vx.geometric_normal(flat, 3, 0, 3), [0, -1, 0])     # bottom: outside the brick counts as empty
    full = np.ones(512, np.uint32)
    assert np.allclose(vx.geometric_normal(full, 7, 7, 7), np.ones(3) / np.sqrt(3), atol=1e-7)  # brick corner
    assert np.allclose(vx.geometric_normal(full, 0, 4, 4), [-1, 0, 0])
    assert np.allclose(vx.geometric_normal(full, 4, 4, 4), [0, 1, 0])      # interior fallback


def test_normal_pool_layout(vx):
    rng = np.random.default_rng(5)
    bricks = (rng.random((3, 512)) < 0.4).astype(np.uint32)
    bricks[2, :16] = 0                                   # an all-empty block
    built = vx.BuiltScene(np.zeros((1, 2), np.uint32), bricks, None, np.zeros((3, 3), np.int32), 0,
                          vx.VxOctreeConfig())
    pool = built.encode_normals().normals
    assert pool.shape == (3, 64)
    assert pool[2, 0] == 0 and pool[2, 1] == 0
    for brick in range(3):
        for blk in (0, 5, 31):
            n = np.array([vx.geometric_normal(bricks[brick], lin & 7, (lin >> 3) & 7, lin >> 6)
                          if bricks[brick, lin] else (0.0, 1.0, 0.0) for lin in range(16 * blk, 16 * blk + 16)],
                         np.float32)
            if not bricks[brick, 16 * blk:16 * blk + 16].any():
                continue
            assert (int(pool[brick, 2 * blk]), int(pool[brick, 2 * blk + 1])) == vx.dxt_normal_encode_block(n)


def test_shader_bc1_decode_matches_cpu_decode(orc):
    rng = np.random.default_rng(11)
    worst = 0.0
    for _ in range(2000):
        blk = int(rng.integers(0, 2 ** 64, dtype=np.uint64))
        for t in (0, 5, 10, 15):
            a, b = orc.bc1_decode_shader(blk, t), orc.bc1_decode(blk, t)
            worst = max(worst, float(np.abs(a.astype(np.float64) - b).max()))
    # the two formulas interpolate the same RGB565 endpoints; the shader keeps the
    # fields' lower bits as fixed point (below 1/32 of a step), so they agree to ~2^-5
    assert worst < 2 ** -4


def _shade_f64(orc, oscene, built, rays, hits, v):
    """float64 restatement of the compressed shader's shading for each pixel."""
    L = np.array([1.0, 1.0, -1.0]) / np.sqrt(3.0)
    out = np.zeros((len(hits), 3))
    d = rays["d"].astype(np.float64)
    du = d / np.linalg.norm(d, axis=1, keepdims=True)
    for i, h in enumerate(hits):
        if not h["flags"] & 1:
            out[i] = np.array([0.5, 0.7, 1.0]) * (1.0 - v[i] * 0.5)
            continue
        blk = int(h["brick"]) * 32 + (int(h["voxel"]) >> 4)
        t = int(h["voxel"]) & 15
        block = int(built.bc1.reshape(-1)[blk])
        head, sel = block & 0xFFFFFFFF, (block >> (32 + 2 * t)) & 3
        w0 = [1.0, 0.0, 2.0 / 3.0, 1.0 / 3.0][sel]
        ref0 = np.array([(head << 16) & 0xFFFFFFFF, (head << 21) & 0xFFFFFFFF, (head << 27) & 0xFFFFFFFF]) / 2.0 ** 32
        ref1 = np.array([head, (head << 5) & 0xFFFFFFFF, (head << 11) & 0xFFFFFFFF]) / 2.0 ** 32
        rgb = w0 * ref0 + (1.0 - w0) * ref1
        last = (int(h["flags"]) >> 4) & 3
        n = np.zeros(3)
        if last == 0:
            a = np.abs(du[i])
            ax = 0 if (a[0] > a[1] and a[0] > a[2]) else (1 if a[1] > a[2] else 2)
            n[ax] = -np.sign(du[i, ax])
        else:
            pool = built.normals.reshape(-1)
            nn = _unit(orc.dxt_normal_decode(int(pool[2 * blk]), int(pool[2 * blk + 1]), t))
            nn[last - 1] += -(1.0 if du[i, last - 1] >= 0 else -1.0) * 0.5
            n = _unit(nn)
        out[i] = rgb * (0.3 + max(n @ L, 0.0) * 0.7)
    return np.clip(np.floor(np.clip(out, 0, 1) * 255 + 0.5), 0, 255).astype(np.int64)


def test_oracle_shading_matches_float64(vx, orc):
    grid, cam = vx.gen_cave(32, 9, 0.5)
    built = vx.build_from_dense(grid, bc1=True).encode_normals()
    oscene = orc.OracleScene.from_built(built)
    W, H = 64, 48
    rgba = orc.render(oscene, cam.as_floats(), W, H, 1e30)
    rays = orc.camera_rays(cam.as_floats(), W, H, 1e30)
    hits, _ = orc.trace(oscene, rays)
    v = ((np.arange(H, dtype=np.float32) + 0.5) / H).repeat(W)
    want = _shade_f64(orc, oscene, built, rays, hits, v)
    got = np.stack([(rgba >> (8 * c)) & 255 for c in range(3)], 1).astype(np.int64)
    assert ((rgba >> 24) == 255).all()
    assert np.abs(got - want).max() <= 1
    assert (got == want).mean() > 0.97
    # the step axis into the hit voxel is the entry-face axis whenever a step happened
    hit = (hits["flags"] & 1) == 1
    last = (hits["flags"] >> 4) & 3
    axis = (hits["flags"] >> 1) & 3
    stepped = hit & (last > 0)
    assert stepped.any() and (hit & (last == 0)).any()
    assert (last[stepped] - 1 == axis[stepped]).all()
