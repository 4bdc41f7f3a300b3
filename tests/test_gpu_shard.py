"""One GPU's share of the low cells (shard sweep) against the oracle's shard semantics."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MODS = [4294967291, 4294966997, 2147483647, 1000003]


@pytest.mark.parametrize("variant", [0, 2, 3])
@pytest.mark.parametrize("n,slices,picks", [(12, 1, [0]), (12, 5, [0, 2, 4]), (16, 16, [3, 15]),
                                            (27, 1 << 30, [7, (1 << 29) + 12345])])
def test_shard_sweep_matches_oracle(gpu_ctx, oracle, n, slices, picks, variant):
    import ofp
    gpu_ctx.set_variant(variant)
    try:
        _check_slices(gpu_ctx, oracle, n, slices, picks)
    finally:
        gpu_ctx.set_variant(0)


def _check_slices(gpu_ctx, oracle, n, slices, picks):
    import ofp
    for s in picks:
        # a slice narrower than a prefix class is empty by construction: take the
        # first non-empty slice at or after the pick
        while ofp._lib.shard_plan(n, len(MODS), slices, s)["blocked_count"] == 0:
            s += 1
        plan = ofp._lib.shard_plan(n, len(MODS), slices, s)
        assert plan["main_count"] + plan["blocked_count"] < 3_000_000
        print(f"case shard n={n} slices={slices} slice={s}", flush=True)
        info, main, blk = gpu_ctx.shard_sweep(n, MODS, slices, s, seed=4242, copy_out=True)
        assert info["stray_stores"] == 0
        assert info["blocked_count"] > 0
        for j, p in enumerate(MODS):
            em, eb = oracle.shard_cells(n, p, j, 4242, info["first_blocked"], info["blocked_count"],
                                        info["first_main"], info["main_count"], info["low_cells"], 1)
            assert np.array_equal(main[:, j], em), (n, s, j)
            assert np.array_equal(blk[:, j], eb), (n, s, j)
