"""Writes tests/golden/golden_small.npz: a seeded 32^3 BC1 cave with its DXT normal
pool, camera and random rays, the oracle's hit records for them, and the
oracle's 64x48 RGBA8 images of the camera view (with and without the normal pool).

The fixture freezes today's oracle output so later changes to the oracle, the
builder or the kernels that alter any bit are caught on every machine. The
scene buffers are stored too, so the fixture does not depend on the generator.
Run from the repo root: python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import __graft_entry__ as entry  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def main() -> None:
    vx = entry.load_package()
    grid, cam = vx.gen_cave(32, 9, 0.5)
    built = vx.build_from_dense(grid, bc1=True).encode_normals()
    rays_cam = orc.camera_rays(cam.as_floats(), 64, 48, 1e30)
    rng = np.random.default_rng(2024)
    n = 3000
    rays_rnd = np.zeros(n, orc.RAY_DTYPE)
    rays_rnd["o"] = rng.uniform(-16, 48, (n, 3))
    rays_rnd["d"] = rng.normal(size=(n, 3))
    rays_rnd["tmin"] = np.where(rng.random(n) < 0.25, rng.uniform(0, 16, n), 0.0)
    rays_rnd["tmax"] = np.where(rng.random(n) < 0.25, rng.uniform(0, 40, n), 1e30)
    rays = np.concatenate([rays_cam, rays_rnd])
    oscene = orc.OracleScene.from_built(built)
    hits, counts = orc.trace(oscene, rays)
    rgba_nrm = orc.render(oscene, cam.as_floats(), 64, 48, 1e30)
    built_face = vx.BuiltScene(built.nodes, built.brick_data, built.bc1, built.brick_coords, built.internal_count,
                               built.config)
    rgba_face = orc.render(orc.OracleScene.from_built(built_face), cam.as_floats(), 64, 48, 1e30)
    c = built.config
    out = Path(__file__).with_name("golden_small.npz")
    np.savez_compressed(
        out, nodes=built.nodes, brick_data=built.brick_data, bc1=built.bc1, brick_coords=built.brick_coords,
        normals=built.normals, rgba_nrm=rgba_nrm, rgba_face=rgba_face,
        grid_sha256=np.frombuffer(hashlib.sha256(grid.tobytes()).digest(), np.uint8),
        config=np.frombuffer(bytes(c), np.uint8), camera=np.frombuffer(bytes(cam), np.uint8),
        rays=rays.view(np.uint8), hits=hits.view(np.uint8), counts=counts)
    print(f"wrote {out} ({out.stat().st_size} bytes): {built.node_count} descriptors, {built.brick_count} bricks, "
          f"{len(rays)} rays, {int((hits['flags'] & 1).sum())} hits")


if __name__ == "__main__":
    main()
