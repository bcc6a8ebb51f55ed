"""vx_bc1_encode_device: the BC1 pool encoded on the GPU equals the host
builder's bit for bit.

* the 256^3 cave's whole brick pool (RGBA8 words) against the pool the host
  builder wrote for it;
* the material-id Cornell room (ids up to 70000) against the builder's
  VX_BC1_FROM_MATERIAL pool, with the palette of MaterialIdToColor over ids
  0-511;
* seeded bricks made to hit the encoder's edge cases against the host encoder
  block by block (vx_bc1_encode_block on the builder's colour conversion,
  channel / 255): empty voxels among solid ones, all-empty blocks (stored as 0),
  one-colour blocks, two-colour blocks with ties between equally distant pairs,
  colours at 0 and 255, random colours;
* a brick pool that does not start 16-byte aligned is refused.
"""
from __future__ import annotations

import numpy as np
import pytest

from test_materials import _material_grid

pytestmark = pytest.mark.gpu


def _host_pool(vx, bricks: np.ndarray) -> np.ndarray:
    """The builder's BC1 pool of RGBA8 bricks [n, 512], restated through the host encoder block by block."""
    out = np.zeros((len(bricks), 32), np.uint64)
    for b in range(len(bricks)):
        for k in range(32):
            w = bricks[b, 16 * k:16 * k + 16]
            if not w.any():
                continue
            cols = np.stack([(w & 255), (w >> 8) & 255, (w >> 16) & 255], axis=1).astype(np.float32) / np.float32(255)
            cols[w == 0] = 0.0
            out[b, k] = vx.bc1_encode_block(cols)
    return out


def test_cave_pool_equals_the_builders(vx, cave):
    import torch
    _, _, built = cave
    bd = torch.from_numpy(built.brick_data.view(np.int32)).cuda()
    got = vx.bc1_encode_device(bd)
    torch.cuda.synchronize()
    assert got.cpu().numpy().view(np.uint64).tobytes() == built.bc1.tobytes()
    assert len(built.bc1) > 10_000


def test_material_ids_with_the_cpu_palette(vx):
    import torch
    ids, _ = _material_grid(vx)
    built = vx.build_from_dense(ids, bc1=True, materials=True)
    pal = torch.from_numpy(vx.material_id_palette()).cuda()
    got = vx.bc1_encode_device(torch.from_numpy(built.brick_data.view(np.int32)).cuda(), palette=pal)
    torch.cuda.synchronize()
    assert got.cpu().numpy().view(np.uint64).tobytes() == built.bc1.tobytes()


def test_edge_case_blocks_against_the_host_encoder(vx):
    import torch
    rng = np.random.default_rng(21)
    n = 96
    bricks = (rng.integers(1, 1 << 24, (n, 512), dtype=np.uint32) | np.uint32(0xFF000000))
    bricks[rng.random((n, 512)) < 0.35] = 0                               # empty voxels among solid ones
    blocks = bricks.reshape(n, 32, 16)
    blocks[0:8] = 0                                                       # all-empty blocks and bricks
    blocks[8:16] = np.uint32(0xFF336699)                                  # one colour
    two = np.where(rng.random((8, 32, 16)) < 0.5, np.uint32(0xFF0000FF), np.uint32(0xFFFF0000))
    blocks[16:24] = two                                                   # two colours: many equally distant pairs
    blocks[24:32] = np.where(rng.random((8, 32, 16)) < 0.5, np.uint32(0xFFFFFFFF), np.uint32(0xFF000000))
    blocks[32:40, :, :] = np.where(rng.random((8, 32, 16)) < 0.5, blocks[32:40], 0)
    bricks = blocks.reshape(n, 512)
    want = _host_pool(vx, bricks)
    got = vx.bc1_encode_device(torch.from_numpy(bricks.view(np.int32)).cuda())
    torch.cuda.synchronize()
    got = got.cpu().numpy().view(np.uint64)
    bad = np.argwhere(got != want)
    assert len(bad) == 0, f"{len(bad)} blocks differ, first (brick, block) {tuple(bad[0])}"
    assert (want[0:8] == 0).all() and (want[8:] != 0).any()


def test_misaligned_pool_is_refused(vx):
    import torch
    bd = torch.zeros(512 * 2 + 1, dtype=torch.int32, device="cuda")
    out = torch.empty((1, 32), dtype=torch.int64, device="cuda")
    st = vx.LIB.vx_bc1_encode_device(bd.data_ptr() + 4, 1, None, out.data_ptr(), None)
    assert st != 0
    assert "aligned" in vx.LIB.vx_last_error().decode()
