"""The seeded cosine-weighted bounce generator (vx_bounce_rays), pinned by what
the distribution fixes, on the bounce rays of the 1080p cave frame (about 2 M
primary hits):
  * every direction is a unit vector in the hemisphere of the hit face's normal;
  * the cosine-weighted density p(w) = cos(theta) / pi gives E[cos theta] = 2/3,
    E[cos^2 theta] = 1/2 and a uniform azimuth about the normal (checked within
    6 standard errors, and by a chi-square test over 36 azimuth bins);
  * every origin lies 1e-3 voxel off the hit voxel's entry face, on its normal's
    side, within the face's extent;
  * a missed primary gets a zero direction; the same seed gives the same rays and
    another seed other rays;
  * a hit record whose face-axis field reads 3 (the trace writes 0-2; a corrupt
    or foreign buffer) is treated as a miss, and valid records beside it are
    unchanged.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import global_voxels

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bounce(vx, cave, gpu_ctx):
    import torch
    grid, cam, built = cave
    scene = vx.Scene(gpu_ctx, built)
    primary = vx.camera_rays(cam, 1920, 1080, 1e30)
    hits = scene.trace(primary)
    out = scene.bounce_rays(primary, hits, seed=7)
    again = scene.bounce_rays(primary, hits, seed=7)
    other = scene.bounce_rays(primary, hits, seed=8)
    torch.cuda.synchronize()
    return (built, primary.cpu().numpy().view(vx.RAY_DTYPE), vx.hits_to_numpy(hits),
            out.cpu().numpy().view(vx.RAY_DTYPE), again.cpu().numpy(), other.cpu().numpy())


def _normals(h):
    n = np.zeros((len(h), 3))
    axis = (h["flags"] >> 1) & 3
    n[np.arange(len(h)), axis] = np.where((h["flags"] >> 3) & 1, 1.0, -1.0)
    return n, axis


def test_directions_are_cosine_distributed_about_the_normal(bounce):
    built, _, hits, b, _, _ = bounce
    hit = (hits["flags"] & 1) == 1
    assert hit.sum() > 1_000_000
    n, axis = _normals(hits[hit])
    d = b["d"][hit].astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1.0).max() < 1e-5
    cos = (d * n).sum(1)
    assert cos.min() >= 0.0                                   # the hemisphere of the face normal
    m = len(cos)
    se1 = np.sqrt(1 / 18 / m)                                 # Var[cos] = 1/2 - 4/9
    se2 = np.sqrt(1 / 12 / m)                                 # Var[cos^2] = 1/3 - 1/4
    assert abs(cos.mean() - 2 / 3) < 6 * se1, cos.mean()
    assert abs((cos ** 2).mean() - 1 / 2) < 6 * se2, (cos ** 2).mean()
    # azimuth about the normal in the face plane (tangent axes (axis+1)%3 and (axis+2)%3)
    t1 = d[np.arange(m), (axis + 1) % 3]
    t2 = d[np.arange(m), (axis + 2) % 3]
    phi = np.arctan2(t2, t1)
    counts, _ = np.histogram(phi, bins=36, range=(-np.pi, np.pi))
    expect = m / 36
    chi2 = ((counts - expect) ** 2 / expect).sum()
    assert chi2 < 90, chi2                                    # 35 degrees of freedom: p ~ 1e-6 at 90
    # the radial distribution: sin^2 theta = 1 - cos^2 is uniform on [0, 1] (Malley's disk)
    counts, _ = np.histogram(1 - cos ** 2, bins=20, range=(0, 1))
    chi2 = ((counts - m / 20) ** 2 / (m / 20)).sum()
    assert chi2 < 65, chi2                                    # 19 degrees of freedom


def test_origins_sit_just_off_the_hit_face(bounce):
    built, _, hits, b, _, _ = bounce
    hit = (hits["flags"] & 1) == 1
    h = hits[hit]
    vox = global_voxels(built, h).astype(np.float64)
    n, axis = _normals(h)
    o = b["o"][hit].astype(np.float64)
    rows = np.arange(len(h))
    plane = vox[rows, axis] + np.where(n[rows, axis] > 0, 1.0, 0.0)   # the entry face's plane
    off = (o[rows, axis] - plane) * n[rows, axis]                       # signed distance along the normal
    assert np.abs(off - 1e-3).max() < 2.5e-4, np.abs(off - 1e-3).max()  # FP32 at coordinates up to 256
    for k in (1, 2):                                                    # inside the face's extent
        ax = (axis + k) % 3
        c = o[rows, ax]
        assert ((c >= vox[rows, ax] - 1e-4) & (c <= vox[rows, ax] + 1 + 1e-4)).all()


def test_misses_and_seeds(bounce):
    _, prim, hits, b, again, other = bounce
    miss = (hits["flags"] & 1) == 0
    assert miss.any()
    assert (b["d"][miss] == 0).all()
    assert np.array_equal(b.view(np.uint8), again)
    assert not np.array_equal(b.view(np.uint8), other)


def test_corrupt_face_axis_counts_as_miss(vx, cave, gpu_ctx, bounce):
    import torch
    grid, cam, built = cave
    _, prim, hits, b, _, _ = bounce
    scene = vx.Scene(gpu_ctx, built)
    bad_hits = hits.copy()
    rng = np.random.default_rng(3)
    corrupt = rng.random(len(bad_hits)) < 0.1
    bad_hits["flags"][corrupt] |= np.uint32(0x7)                 # hit bit set, axis field 3
    out = scene.bounce_rays(torch.from_numpy(prim.view(np.uint8).reshape(-1).copy()).cuda(),
                            torch.from_numpy(bad_hits.view(np.uint8).reshape(-1).copy()).cuda(), seed=7)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(vx.RAY_DTYPE)
    assert (got["d"][corrupt] == 0).all()
    assert np.array_equal(got["o"][corrupt], prim["o"][corrupt])
    assert np.array_equal(got[~corrupt].view(np.uint8), b[~corrupt].view(np.uint8))
    scene.close()
