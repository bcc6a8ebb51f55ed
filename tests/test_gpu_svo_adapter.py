"""The C++ ISVOStructure-shaped adapter on the GPU, driven from a C++ host with no
Python in its process (tests/svo_consumer trace).

On the seeded 128^3 cave with BC1 colours the host casts rays one per call
through castRay, castRayLOD (cone growth 0.01), castRayScreenSpaceLOD (fovY 1.2,
90 lines) and castRayWithLOD (fromCamera(1.0, 1080) with bias 10), then 200,000
rays through castRays (the first 3000 being castRay's) and castRayWithLOD's rays
again through castRays. Every RayHit must equal, byte for byte, the reference's
RayHit built from the oracle's record of the same ray (tests/svo_expect.py), and
the batch casts must equal the one-ray casts.
"""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

import svo_expect as se

pytestmark = pytest.mark.gpu


def test_cpp_host_casts_rays_bit_exact(vx, orc, gpu_ctx, tmp_path):
    if not se.EXE.exists():
        pytest.fail("tests/svo_consumer is not built (run __graft_entry__.build())")
    out = tmp_path / "trace.bin"
    r = subprocess.run([str(se.EXE), "trace", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    groups = se.read_groups(out.read_bytes(), orc.RAY_DTYPE)
    assert [len(g[0]) for g in groups] == [3000, 1000, 1000, 1000, 200000, 1000]

    grid, _ = vx.gen_cave(128, 5, 0.5)
    built = vx.build_from_dense(grid, bc1=True)
    scene = orc.OracleScene.from_built(built)
    lods = [None, vx.VxLod(0.0, 0.01), vx.lod_from_camera(1.2, 90),
            vx.lod_with_bias(vx.lod_from_camera(1.0, 1080), 10.0), None, None]
    lods[5] = lods[3]
    names = ["castRay", "castRayLOD", "castRayScreenSpaceLOD", "castRayWithLOD", "castRays", "castRays(lod)"]
    for (rays, got), lod, name in zip(groups, lods, names):
        rays = np.ascontiguousarray(rays)
        if lod is None:
            want_hits, _ = orc.trace(scene, rays, nthreads=8)
        else:
            want_hits = orc.trace_lod(scene, rays, lod.ray_orig_size, lod.ray_dir_size, nthreads=8)
        want = se.expected_rayhits(rays, want_hits, 7, 128.0)
        bad = se.rec_diff(got, want)
        assert len(bad) == 0, f"{name}: {len(bad)} of {len(rays)} RayHits differ, first {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"
        assert (got["hit"] == 1).sum() > len(rays) // 4, name
    assert (groups[3][1]["lod"] == 1).sum() > 50  # the biased cone stops some rays at coarse children
    # one ray per call and many per call give the same records
    assert se.rec_diff(groups[4][1][:3000], groups[0][1]).size == 0
    assert se.rec_diff(groups[5][1], groups[3][1]).size == 0
