"""An independent restatement of VIXEN's VSCA v1 scene-cache layout, for the tests.

Field by field from the reference's declarations, packed with struct:
  * the file and entry order of VoxelSceneCacher::SerializeToFile
    (libraries/CashSystem/src/VoxelSceneCacher.cpp:210-258);
  * OctreeConfig as libraries/CashSystem/include/VoxelSceneCacher.h:100-133
    declares it: 8 int32, gridMin/gridMax as vec3 + pad, two column-major mat4,
    16 floats of padding, then the non-UBO float worldGridSize at offset 256,
    so sizeof(OctreeConfig) = 260 and the file holds all 260 bytes (:254, :336);
  * the values VoxelSceneCacher::BuildOctree fills in (VoxelSceneCacher.cpp:592-625).
"""
from __future__ import annotations

import math
import struct

import numpy as np

MAGIC = 0x56534341
CONFIG_BYTES = 260  # sizeof(OctreeConfig): 256-byte UBO + worldGridSize


def key(scene_type, resolution, density, seed):
    dq = int(np.float32(density) * np.float32(100.0))  # static_cast<uint32_t>(density * 100.0f)
    h = scene_type
    for v in (resolution, dq, seed):
        h = (h * 31 + v) % 2 ** 64
    return h


def config_fields(resolution: int, world_grid_size: float = 10.0) -> dict:
    """OctreeConfig as BuildOctree fills it for a resolution^3 grid of 8^3 bricks."""
    L = int(math.ceil(math.log2(resolution)))             # resolutionLevels = maxLevels
    brick_user_scale = L - 3
    ltw = np.diag([world_grid_size, world_grid_size, world_grid_size, 1.0]).astype(np.float32)
    wtl = np.diag([np.float32(1.0) / np.float32(world_grid_size)] * 3 + [1.0]).astype(np.float32)
    return dict(esvoMaxScale=22, userMaxLevels=L, brickDepthLevels=3, brickSize=8,
                minESVOScale=22 - L + 1, brickESVOScale=22 - (L - 1 - brick_user_scale),
                bricksPerAxis=resolution // 8, gridMin=(0.0, 0.0, 0.0), gridMax=(float(resolution),) * 3,
                localToWorld=ltw, worldToLocal=wtl, worldGridSize=world_grid_size)


def pack_config(f: dict) -> bytes:
    """The 260 bytes SerializeToFile writes for one OctreeConfig."""
    out = struct.pack("<8i", f["esvoMaxScale"], f["userMaxLevels"], f["brickDepthLevels"], f["brickSize"],
                      f["minESVOScale"], f["brickESVOScale"], f["bricksPerAxis"], 0)
    out += struct.pack("<3ff", *f["gridMin"], 0.0) + struct.pack("<3ff", *f["gridMax"], 0.0)
    # glm::mat4 is column-major: column c holds elements [c][0..3]
    out += np.ascontiguousarray(np.asarray(f["localToWorld"], np.float32).T).tobytes()
    out += np.ascontiguousarray(np.asarray(f["worldToLocal"], np.float32).T).tobytes()
    out += bytes(64)                                      # _padding4[16]
    assert len(out) == 256
    out += struct.pack("<f", f["worldGridSize"])          # offset 256
    assert len(out) == CONFIG_BYTES
    return out


def vec(data: bytes, elem: int = 1) -> bytes:
    """WriteVector: u64 element count, then the elements."""
    return struct.pack("<Q", len(data) // elem) + data


def entry(scene_type, resolution, density, seed, nodes: bytes, bricks: bytes, materials: bytes, colors: bytes,
          normals: bytes, lookup: bytes, config: bytes, node_count, brick_count, solid) -> bytes:
    return (struct.pack("<Q", key(scene_type, resolution, density, seed)) +
            struct.pack("<IIfI", scene_type, resolution, density, seed) +
            vec(nodes) + vec(bricks) + vec(materials, 32) + vec(colors) + vec(normals) + vec(lookup, 4) + config +
            struct.pack("<IIIII", node_count, brick_count, solid, resolution, scene_type))


def file(entries: list[bytes], version=1, magic=MAGIC) -> bytes:
    return struct.pack("<III", magic, version, len(entries)) + b"".join(entries)


def built_entry(vx, built, scene_type, resolution, density, seed, materials=b"", solid=None,
                world_grid_size=10.0, config: bytes | None = None) -> bytes:
    """The bytes of one entry holding a scene built here, with its config restated from the reference's
    formulas (not taken from the builder), unless config is given."""
    lookup = vx.brick_grid_lookup(built)
    solid = int((built.brick_data != 0).sum()) if solid is None else solid
    cfg = config if config is not None else pack_config(config_fields(resolution, world_grid_size))
    return entry(scene_type, resolution, density, seed, built.nodes.tobytes(), built.brick_data.tobytes(), materials,
                 b"" if built.bc1 is None else built.bc1.tobytes(),
                 b"" if built.normals is None else built.normals.tobytes(), lookup.tobytes(), cfg,
                 built.node_count, built.brick_count, solid)
