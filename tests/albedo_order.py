"""The deterministic albedo reduction's summation order, restated in numpy
float32 for the tests, and an adversarial layout of per-pixel rows.

vx_albedo_backward (deterministic mode) sums each row's contributions
dL/dpixel * shade in this order:
  * a row hit by at most 32 pixels: sequentially, in pixel order;
  * a row hit by 33 to 4096 pixels: 64 lanes, lane t adding the row's elements
    t, t+128, ... (in sorted, that is pixel, order) into one partial sum and
    t+64, t+192, ... into another, then s0 + s1 per lane and a halving tree over
    the lanes (h = 32, 16, ..., 1);
  * a longer row: the same with 256 threads (elements t, t+512, ... and t+256,
    t+768, ...; tree h = 128, 64, ..., 1);
and adds each row's sum to the row's gradient. Misses carry row 0xFFFFFFFF.
"""
from __future__ import annotations

import numpy as np

MISS = 0xFFFFFFFF
SHORT = 32
WAVE_RUN = 4096  # the longest row summed by one wave of 64 lanes; longer rows use 256 threads


def _tree_sum(v: np.ndarray, width: int) -> np.ndarray:
    """Elements of v [n, 3] summed by `width` lanes: lane t adds t, t+2w, ... and t+w, t+3w, ..., then s0 + s1
    and a halving tree over the lanes."""
    n = len(v)
    lane = np.arange(width)
    s0 = np.zeros((width, 3), np.float32)
    s1 = np.zeros((width, 3), np.float32)
    for k0 in range(0, n, 2 * width):
        i0, i1 = k0 + lane, k0 + width + lane
        m0, m1 = i0 < n, i1 < n
        s0[m0] = s0[m0] + v[i0[m0]]
        s1[m1] = s1[m1] + v[i1[m1]]
    part = s0 + s1
    h = width // 2
    while h > 0:
        part[:h] = part[:h] + part[h:2 * h]
        h //= 2
    return part[0]


def restated_gradient(rows: np.ndarray, shade: np.ndarray, gimg: np.ndarray, grad0: np.ndarray) -> np.ndarray:
    c = gimg.reshape(-1, 3) * shade[:, None]  # float32 products, as the adjoint kernel forms them
    p = np.flatnonzero(rows != MISS)
    ps = p[np.lexsort((p, rows[p]))]  # by row, then pixel
    rs = rows[ps]
    if len(rs) == 0:
        return grad0.copy()
    starts = np.flatnonzero(np.r_[True, rs[1:] != rs[:-1]])
    ends = np.r_[starts[1:], len(rs)]
    lens = ends - starts
    g = grad0.copy()
    # short rows, all at once: element k of every row longer than k, in order
    sh = lens <= SHORT
    sb, sl = starts[sh], lens[sh]
    s = np.zeros((len(sb), 3), np.float32)
    for k in range(SHORT):
        m = sl > k
        s[m] = s[m] + c[ps[sb[m] + k]]
    g[rs[sb]] = g[rs[sb]] + s
    # long rows: the 64-lane order up to WAVE_RUN pixels, the 256-thread order beyond
    for b, n in zip(starts[~sh], lens[~sh]):
        g[rs[b]] = g[rs[b]] + _tree_sum(c[ps[b:b + n]], 64 if n <= WAVE_RUN else 256)
    return g


def adversarial_rows(rng, px: int, V: int) -> np.ndarray:
    rows = np.full(px, MISS, np.uint32)
    order = rng.permutation(px)
    pos = 0

    def put(row, count):
        nonlocal pos
        rows[order[pos:pos + count]] = row
        pos += count

    picks = rng.choice(V - 2, size=2000, replace=False) + 1
    put(int(picks[0]), min(100_000, px // 4))  # one row over many windows
    for r in picks[1:201]:
        put(int(r), 33)  # the shortest long rows
    for r in picks[201:401]:
        put(int(r), 32)  # the longest short rows
    for r in picks[401:451]:
        put(int(r), int(rng.integers(500, 700)))  # long rows across windows
    put(int(picks[451]), 4096)  # the longest row one wave sums
    put(int(picks[452]), 4097)  # the shortest row the whole workgroup sums
    put(0, 7)
    put(V - 1, 40)
    # the rest: rows hit by 1-5 pixels, leaving about a tenth of the image as misses
    while pos < px - px // 10:
        put(int(rng.integers(0, V)), int(min(rng.integers(1, 6), px - px // 10 - pos)))
    return rows
