"""What the C++ ISVOStructure adapter must return, restated in numpy from the
reference: a GPU hit record becomes ISVOStructure::RayHit as
libraries/SVO/src/SVOBrickDDA.cpp:505-511 builds a voxel hit (tMax = tMin + the
voxel's edge, hitPoint = origin + normalize(direction) * tMin, the entry face's
normal, scale = maxLevels - 1), as libraries/SVO/src/SVOTraversal.cpp:637-647
builds an LOD stop (normal (0,1,0), tMax = tMin + esvoScaleToWorldSize,
libraries/SVO/include/SVOLOD.h:189-194, scale = the user level) and as
SVOTraversal.cpp:524-526 builds a miss (value-initialised: all zero). Also the
file layout tests/svo_consumer writes.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

EXE = Path(__file__).resolve().parent / "svo_consumer"

REC_DTYPE = np.dtype([("entity", "<u8"), ("hitPoint", "<f4", 3), ("normal", "<f4", 3), ("tMin", "<f4"),
                      ("tMax", "<f4"), ("scale", "<i4"), ("hit", "<u4"), ("brick", "<u4"), ("voxel", "<u4"),
                      ("leafNode", "<u4"), ("color", "<f4", 3), ("iterations", "<u4"), ("lod", "<u4")])
assert REC_DTYPE.itemsize == 80


def read_groups(raw: bytes, ray_dtype) -> list[tuple[np.ndarray, np.ndarray]]:
    """The (rays, records) groups of a svo_consumer output file."""
    out, at = [], 0
    while at < len(raw):
        n = int(np.frombuffer(raw, "<u8", 1, at)[0])
        at += 8
        rays = np.frombuffer(raw, ray_dtype, n, at)
        at += n * ray_dtype.itemsize
        recs = np.frombuffer(raw, REC_DTYPE, n, at)
        at += n * REC_DTYPE.itemsize
        out.append((rays, recs))
    return out


def expected_rayhits(rays: np.ndarray, hits: np.ndarray, max_levels: int, world_x: float) -> np.ndarray:
    """RayHit records the reference would build from these hit records (float32 arithmetic throughout)."""
    f32 = np.float32
    n = len(rays)
    out = np.zeros(n, REC_DTYPE)
    flags = hits["flags"].astype(np.uint32)
    out["brick"], out["voxel"], out["leafNode"] = hits["brick"], hits["voxel"], hits["leaf_node"]
    out["iterations"] = flags >> 16
    hit = (flags & 1) != 0
    lod = hit & ((flags & 64) != 0)
    d = rays["d"].astype(f32)
    with np.errstate(all="ignore"):
        len2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]  # glm::dot: x, y, then z
        inv = f32(1.0) / np.sqrt(len2)
        dn = d * inv[:, None]                                              # glm::normalize
        t = hits["t"].astype(f32)
        hp = rays["o"].astype(f32) + dn * t[:, None]
    out["hit"] = hit
    out["tMin"] = np.where(hit, t, f32(0))
    out["hitPoint"] = np.where(hit[:, None], hp, f32(0))
    out["color"] = np.where(hit[:, None], hits["rgb"], f32(0))
    ext = f32(world_x)
    # voxel hits
    vox = hit & ~lod
    axis = (flags >> 1) & 3
    sign = np.where((flags & 8) != 0, f32(1), f32(-1))
    normal = np.zeros((n, 3), f32)
    for a in range(3):
        normal[:, a] = np.where(vox & (axis == a), sign, f32(0))
    normal[:, 1] = np.where(lod, f32(1), normal[:, 1])
    out["normal"] = normal
    lod_level = ((flags >> 8) & 31).astype(np.int32)
    out["scale"] = np.where(lod, lod_level, np.where(vox, max_levels - 1, 0))
    edge_vox = ext / f32(2 ** max_levels)
    edge_lod = np.ldexp(f32(1), lod_level - max_levels).astype(f32) * ext
    with np.errstate(all="ignore"):
        out["tMax"] = np.where(lod, t + edge_lod, np.where(vox, t + edge_vox, f32(0)))
    out["lod"] = lod
    return out


def rec_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Indices of records whose 80 bytes differ."""
    wa = a.view(np.uint32).reshape(-1, 20)
    wb = b.view(np.uint32).reshape(-1, 20)
    return np.nonzero((wa != wb).any(1))[0]
