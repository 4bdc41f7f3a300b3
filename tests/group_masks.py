"""Occupancy masks of every word in the global rank order, computed with numpy
from the oracle's height table (test helper: the window order of occupancy
groups is the global order sorted stably by mask)."""
from __future__ import annotations

import numpy as np


def height_table(rows: int, cols: int = 34) -> np.ndarray:
    d = np.zeros((rows, cols), dtype=np.int64)
    d[0, 0] = 1
    for w in range(1, rows):
        d[w, :] = d[w - 1, :]
        d[w, 1:] += d[w - 1, :-1]
        d[w, :-1] += d[w - 1, 1:]
    return d


def occupancy_masks(count: int, width: int, low: int, lo: int = 0) -> np.ndarray:
    """mask[i] has bit pos-lo set, for lo <= pos < lo+low, where word i (width
    `width`, lexicographic rank N < R < L from the top, entered at height 1) is
    occupied at pos."""
    d = height_table(width + 2)
    r = np.arange(count, dtype=np.int64)
    h = np.ones(count, dtype=np.int64)
    mask = np.zeros(count, dtype=np.int64)
    for pos in range(width - 1, -1, -1):
        zn = d[pos, h]
        is_n = r < zn
        r = np.where(is_n, r, r - zn)
        zr = np.where(h > 0, d[pos, np.maximum(h - 1, 0)], 0)
        is_r = ~is_n & (r < zr)
        is_l = ~is_n & ~is_r
        r = np.where(is_l, r - zr, r)
        h = h + is_l.astype(np.int64) - is_r.astype(np.int64)
        if lo <= pos < lo + low:
            mask |= (~is_n).astype(np.int64) << (pos - lo)
    assert (r == 0).all() and (h == 0).all()
    return mask


def window_order(mask: np.ndarray, first_group: int, groups: int) -> np.ndarray:
    """Global indices of the window's words in window order."""
    sel = np.nonzero((mask >= first_group) & (mask < first_group + groups))[0]
    return sel[np.argsort(mask[sel], kind="stable")]


def group_masks(n: int, low: int, high: bool):
    """(main masks, blocked masks) of every word for LOW (positions 0..low-1)
    or HIGH (main positions low+1..n, blocked positions low..n-1) groups."""
    nm, nb = _counts(n)
    if not high:
        return occupancy_masks(nm, n + 1, low), occupancy_masks(nb, n, low)
    fixed = n - low
    return occupancy_masks(nm, n + 1, fixed, low + 1), occupancy_masks(nb, n, fixed, low)


def _counts(n: int):
    d = height_table(n + 3)
    return int(d[n + 1, 1]), int(d[n, 1])
