"""CPU check of the restated deterministic albedo order (tests/albedo_order.py):
its per-row sums must agree with float64 sums of the same contributions within
the float32 rounding bound of that order, on the adversarial row layout the GPU
test uses (rows of 32 and 33 pixels, long rows, one row over many windows,
misses). A dropped, repeated or misplaced element fails it."""
from __future__ import annotations

import numpy as np

from albedo_order import MISS, adversarial_rows, restated_gradient


def test_restatement_matches_float64_sums():
    rng = np.random.default_rng(3)
    V, w, h = 5000, 300, 200
    rows = adversarial_rows(rng, w * h, V)
    shade = rng.random(w * h, dtype=np.float32)
    gimg = rng.standard_normal((h, w, 3), dtype=np.float32)
    grad0 = np.zeros((V, 3), np.float32)
    got = restated_gradient(rows, shade, gimg, grad0).astype(np.float64)
    c = gimg.reshape(-1, 3).astype(np.float64) * shade[:, None].astype(np.float32).astype(np.float64)
    ref = np.zeros((V, 3))
    hit = rows != MISS
    np.add.at(ref, rows[hit].astype(np.int64), c[hit])
    mag = np.zeros((V, 3))
    np.add.at(mag, rows[hit].astype(np.int64), np.abs(c[hit]))
    assert np.all(np.abs(got - ref) <= 1e-5 * mag + 1e-30)
