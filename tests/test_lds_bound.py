"""The 64-bit cell kernels keep only heights 0..21 of the height table in LDS
(LdsCols in oneesan-gpu_amd/csrc/ofp_cell_impl.h).  That is safe when no word of
W <= 28 positions (entered at height 1) is ever above height 19 at a point where it
can still be completed, since the scans read at most two heights above a word's
own.  Checked here with the oracle's height DP (oracle/gridfp_oracle.c or_dp)."""
from __future__ import annotations

LDS_COLS_64 = 22  # LdsCols<uint64_t>::value
MAX_W = 28        # kMaxW


def _max_height(lib, W: int) -> int:
    mx = 0
    for j in range(W + 1):          # symbols read from the top
        for h in range(0, 32):
            if h <= 1 + j and lib.or_dp(W - j, h) > 0:
                mx = max(mx, h)
    return mx


def test_reachable_heights_fit_the_lds_tables(oracle):
    lib = oracle.load()
    worst = 0
    for W in range(2, MAX_W + 1):
        mx = _max_height(lib, W)
        assert mx <= (W + 2) // 2, W  # a word must close every bracket it opens
        worst = max(worst, mx)
    assert worst + 2 < LDS_COLS_64
    assert worst == 14  # W = 27, 28

