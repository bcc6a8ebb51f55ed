"""The ABI driven from C, with no Python or PyTorch in the process.

tests/abi_consumer (built from tests/abi_consumer.c by the package build)
builds the seeded 64^3 Cornell room with BC1 colours, traces its 512x512
camera rays with vx_trace and renders the frame with vx_render. Its trace hits
and render hits must equal the oracle's byte for byte, and its RGBA8 image the
oracle's render word for word.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

from conftest import record_diff

pytestmark = pytest.mark.gpu
EXE = Path(__file__).resolve().parent / "abi_consumer"


def test_c_host_traces_and_renders_bit_exact(vx, orc, gpu_ctx, tmp_path):
    if not EXE.exists():
        pytest.fail("tests/abi_consumer is not built (run __graft_entry__.build())")
    out = tmp_path / "frame.bin"
    r = subprocess.run([str(EXE), str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    n = 512 * 512
    raw = out.read_bytes()
    assert len(raw) == n * 32 * 2 + n * 4
    trace_hits = np.frombuffer(raw[: n * 32], vx.HIT_DTYPE)
    render_hits = np.frombuffer(raw[n * 32: 2 * n * 32], vx.HIT_DTYPE)
    image = np.frombuffer(raw[2 * n * 32:], np.uint32)

    grid, cam = vx.gen_cornell(64, 1)
    built = vx.build_from_dense(grid, bc1=True)
    scene = orc.OracleScene.from_built(built)
    rays = orc.camera_rays(cam.as_floats(), 512, 512, 1e4)
    want, _ = orc.trace(scene, rays, nthreads=8)
    for label, got in (("vx_trace", trace_hits), ("vx_render", render_hits)):
        bad = record_diff(got, want)
        assert len(bad) == 0, f"{label}: {len(bad)} records differ from the oracle; first at {bad[0]}"
    assert np.array_equal(image, orc.render(scene, cam.as_floats(), 512, 512, 1e4))
    assert (want["flags"] & 1).mean() > 0.99
