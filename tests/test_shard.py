"""Shard planning and the oracle's shard semantics (CPU only)."""
from __future__ import annotations

import numpy as np
import pytest

import ofp


@pytest.mark.parametrize("n,slices", [(12, 1), (12, 5), (17, 8), (27, 64), (27, 8)])
def test_slices_partition_by_prefix_class(oracle, n, slices):
    lib = oracle.load()
    W = n + 1
    nm, nb = ofp.state_counts(n)
    plans = [ofp._lib.shard_plan(n, 4, slices, s) for s in range(slices)]
    assert plans[0]["first_blocked"] == 0 and plans[0]["first_main"] == 0
    for a, b in zip(plans, plans[1:]):
        assert a["first_blocked"] + a["blocked_count"] == b["first_blocked"]
        assert a["first_main"] + a["main_count"] == b["first_main"]
    assert plans[-1]["first_blocked"] + plans[-1]["blocked_count"] == nb
    assert plans[-1]["first_main"] + plans[-1]["main_count"] == nm
    lc = plans[0]["low_cells"]
    assert lc == W // 2
    mask_b = ~((1 << (2 * lc)) - 1) & ((1 << 64) - 1)
    for pl in plans[1:]:
        b0, m0 = pl["first_blocked"], pl["first_main"]
        if pl["blocked_count"] == 0:
            continue
        wb, wb_prev = lib.or_unrank(b0, W - 1), lib.or_unrank(b0 - 1, W - 1)
        wm, wm_prev = lib.or_unrank(m0, W), lib.or_unrank(m0 - 1, W)
        # a boundary starts a new prefix class, in both numberings, with the same prefix
        assert wb & mask_b != wb_prev & mask_b
        assert (wm >> 2) & mask_b != (wm_prev >> 2) & mask_b
        assert (wm >> 2) & mask_b == wb & mask_b


@pytest.mark.parametrize("n", [7, 10])
def test_full_window_shard_equals_plain_cells(oracle, n):
    """With one slice the window is the whole state: the shard semantics reduce to
    the plain recurrence on the synthetic inputs."""
    lib = oracle.load()
    nm, nb = ofp.state_counts(n)
    W = n + 1
    lc = W // 2
    seed, mod = 1234, 4294967291
    mw, bw = oracle.shard_cells(n, mod, 1, seed, 0, nb, 0, nm, lc, 1)
    main = np.array([lib.or_remote_residue(seed, i, 1, mod) for i in range(nm)], np.uint32)
    blk = np.array([lib.or_remote_residue(seed, (1 << 62) + b, 1, mod) for b in range(nb)], np.uint32)
    oracle.apply_cells(n, mod, main, blk, lc, 1, True)
    assert np.array_equal(mw, main)
    # after cell 1 the blocked words are untouched since cell 2 in the shard semantics
    m2 = np.array([lib.or_remote_residue(seed, i, 1, mod) for i in range(nm)], np.uint32)
    b2 = np.array([lib.or_remote_residue(seed, (1 << 62) + b, 1, mod) for b in range(nb)], np.uint32)
    oracle.apply_cells(n, mod, m2, b2, lc, 2, True)
    assert np.array_equal(bw, b2)
