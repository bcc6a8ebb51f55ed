"""The automatic store-or-re-trace choice of the differentiable render (host logic, no GPU)."""
from __future__ import annotations

import pytest


def test_stash_bytes(vx):
    # the forward keeps one u32 albedo row and one f32 shade per pixel
    assert vx.HIT_STASH_BYTES_PER_PIXEL == 8


def test_stores_while_the_stash_fits(vx):
    w, h = 1280, 720
    stash = 8 * w * h
    assert vx.stores_hits(w, h, 4 * stash)  # exactly a quarter of free memory
    assert not vx.stores_hits(w, h, 4 * stash - 1)  # memory is tight: re-trace
    assert vx.stores_hits(w, h, stash, memory_fraction=1.0)
    assert not vx.stores_hits(w, h, 10**15, memory_fraction=0.0)
    assert vx.stores_hits(0, 0, 0)  # nothing to keep


@pytest.mark.parametrize("fraction", [-0.1, 1.5])
def test_fraction_out_of_range(vx, fraction):
    with pytest.raises(ValueError):
        vx.stores_hits(16, 16, 1 << 30, memory_fraction=fraction)
