"""Height caps of the first and last rows (libofp's ofp_height_cap) against the oracle.

The product skips, at a capped cell, every blocked word b whose largest height
H(b) exceeds the cap: b's thread does not run, so the main words it owns (NN
and LR for x = N, NR for x = R, NL for x = L, also RN at p = 1) and blocked[b]
keep their values.  Here the oracle's own cell step is applied and then exactly
those words are put back, cell by cell, over whole solves; the answer must be
the oracle's unskipped answer for every modulus.  The forward half of the
claim is also checked directly: before a capped cell of the first rows, every
nonzero word has H <= the cap.
"""
from __future__ import annotations

import numpy as np
import pytest

MODS = [4294967291, 4294966997]


def _heights(lib, count, width):
    out = np.empty(count, np.int32)
    for i in range(count):
        w = lib.or_unrank(i, width)
        h = H = 1
        for pos in range(width - 1, -1, -1):
            v = (w >> (2 * pos)) & 3
            h += (v == 2) - (v == 1)
            H = max(H, h)
        out[i] = H
    return out


def _owner_heights(lib, n, p):
    """For every main word: H of the blocked word whose thread writes it at cell p, or -1 if
    no thread writes it (closure sources and words read only)."""
    W = n + 1
    nm = lib.or_main_count(n)
    own = np.full(nm, -1, np.int32)
    for i in range(nm):
        w = lib.or_unrank(i, W)
        a = (w >> (2 * p)) & 3
        c = (w >> (2 * (p - 1))) & 3
        if a == 0 and c in (0, 1, 2):
            x = c                       # NN -> N, NR -> R, NL -> L
        elif (a, c) == (2, 1):
            x = 0                       # LR is written by the x = N thread
        elif (a, c) == (1, 0) and p == 1:
            x = 1                       # RN at p = 1 (x = R)
        else:
            continue
        low = w & ((1 << (2 * (p - 1))) - 1)
        high = w >> (2 * (p + 1))
        b = (high << (2 * p)) | (x << (2 * (p - 1))) | low
        h = H = 1
        for pos in range(W - 2, -1, -1):
            v = (b >> (2 * pos)) & 3
            h += (v == 2) - (v == 1)
            H = max(H, h)
        own[i] = H
    return own


@pytest.mark.parametrize("n", [3, 4, 5, 6, 7, 8, 9, 10, 11])
def test_capped_sweep_equals_oracle(oracle, n):
    import ofp
    lib = oracle.load()
    W = n + 1
    nm, nb = lib.or_main_count(n), lib.or_blocked_count(n)
    hb = _heights(lib, nb, W - 1)
    hm = _heights(lib, nm, W)
    owners = {p: _owner_heights(lib, n, p) for p in range(1, W)}
    capped_cells = 0
    for mod in MODS:
        main = np.zeros(nm, np.uint32)
        blk = np.zeros(nb, np.uint32)
        main[lib.or_rank(1 << (2 * (W - 1)), W)] = 1
        for r in range(W):
            for p in range(W - 1, 0, -1):
                cap, swept = ofp.height_cap(n, r, p, 100)
                if cap == 0:
                    oracle.apply_cells(n, mod, main, blk, p, p, p != W - 1)
                    continue
                capped_cells += 1
                assert swept == int((hb <= cap).sum())
                if r < W - 2:
                    # forward bound: nothing above the cap is nonzero yet
                    assert not main[hm > cap].any() and not (p != W - 1 and blk[hb > cap].any()), (n, r, p)
                m0, b0 = main.copy(), blk.copy()
                oracle.apply_cells(n, mod, main, blk, p, p, p != W - 1)
                skip_m = owners[p] > cap
                main[skip_m] = m0[skip_m]
                if p >= 2:
                    skip_b = hb > cap
                    blk[skip_b] = b0[skip_b]
        assert int(main[0]) == oracle.solve(n, mod), (n, mod)
    if n >= 3:
        assert capped_cells > 0


def test_cap_schedule_shape():
    """Row 0 and the last row sweep only H <= 1; the middle rows are uncapped."""
    import ofp
    for n in (12, 18, 22):
        W = n + 1
        assert all(ofp.height_cap(n, 0, p)[0] == 1 for p in range(1, W))
        assert all(ofp.height_cap(n, W - 1, p)[0] == 1 for p in range(1, W))
        assert all(ofp.height_cap(n, W // 2, p)[0] == 0 for p in range(1, W))
        assert ofp.height_cap(n, -1, 1)[0] == 0 and ofp.height_cap(n, W, 1)[0] == 0
