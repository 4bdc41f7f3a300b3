"""Group sweeps on the GPU through the C ABI: the high cells W-1..low+1 of a
row run over occupancy groups of the low positions, each group stored alone
with 32-bit ranks local to the group.  Checked bit-exact against the oracle's
full-state cells, against the full-state GPU path, and, at the n = 27 LOW14
shape, against the same groups swept in other windows."""
from __future__ import annotations

import numpy as np
import pytest

import ofp
from group_masks import occupancy_masks, window_order
from ofp.path_bound import PRIMES

pytestmark = pytest.mark.gpu


def _inputs(nm: int, mods, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)


@pytest.mark.parametrize("n,low,k", [(9, 4, 3), (12, 3, 1), (12, 6, 3), (13, 11, 2), (13, 6, 8), (11, 5, 14)])
def test_group_sweep_equals_oracle(gpu_ctx, oracle, n, low, k):
    mods = ([2] + list(PRIMES))[:k] if k > 1 else [PRIMES[0]]
    nm, nb = ofp.state_counts(n)
    W = n + 1
    main = _inputs(nm, mods, 1000 * n + low)
    om = window_order(occupancy_masks(nm, W, low), 0, 1 << low)
    ob = window_order(occupancy_masks(nb, n, low), 0, 1 << low)
    print(f"case n={n} low={low} k={k} all groups", flush=True)
    info, mo, bo = gpu_ctx.group_sweep(n, mods, 0, 1 << low, low, main_in=main[om], blocked_out=True)
    assert info["cells"] == n - low and len(info["cell_ms"]) == n - low and info["launches"] > 0
    for j, m in enumerate(mods):
        em = np.ascontiguousarray(main[:, j])
        eb = np.zeros(nb, np.uint32)
        oracle.apply_cells(n, m, em, eb, W - 1, low + 1, False, pull=True)
        assert np.array_equal(mo[:, j], em[om]), (j, m)
        assert np.array_equal(bo[:, j], eb[ob]), (j, m)


def test_window_of_groups_equals_whole(gpu_ctx):
    """A window of consecutive groups gives its groups' part of the all-groups run."""
    n, low, mods = 13, 6, [PRIMES[0], PRIMES[1]]
    nm, nb = ofp.state_counts(n)
    mm, mb = occupancy_masks(nm, n + 1, low), occupancy_masks(nb, n, low)
    main = _inputs(nm, mods, 7)
    om = window_order(mm, 0, 1 << low)
    _, whole, wblk = gpu_ctx.group_sweep(n, mods, 0, 1 << low, low, main_in=main[om], blocked_out=True)
    out = np.zeros_like(main)
    out[om] = whole
    blk = np.zeros((nb, len(mods)), np.uint32)
    blk[window_order(mb, 0, 1 << low)] = wblk
    for g0, g in ((0, 1), (10, 20), (40, 24)):
        print(f"case window {g0}+{g}", flush=True)
        ow, obw = window_order(mm, g0, g), window_order(mb, g0, g)
        _, mo, bo = gpu_ctx.group_sweep(n, mods, g0, g, low, main_in=main[ow], blocked_out=True)
        assert np.array_equal(mo, out[ow]) and np.array_equal(bo, blk[obw])


def test_group_sweep_equals_full_state_gpu(gpu_ctx):
    """n = 16, LOW8, 2 moduli: the group sweep equals ofp_apply_cells on the
    whole state (which runs its cells as pair passes)."""
    n, low, mods = 16, 8, [PRIMES[0], PRIMES[5]]
    nm, nb = ofp.state_counts(n)
    main = _inputs(nm, mods, 16)
    om = window_order(occupancy_masks(nm, n + 1, low), 0, 1 << low)
    ob = window_order(occupancy_masks(nb, n, low), 0, 1 << low)
    print("case n=16 low=8 group sweep", flush=True)
    _, mo, bo = gpu_ctx.group_sweep(n, mods, 0, 1 << low, low, main_in=main[om], blocked_out=True)
    em, eb = main.copy(), np.zeros((nb, 2), np.uint32)
    print("case n=16 full-state cells 16..9", flush=True)
    gpu_ctx.apply_cells(n, mods, em, eb, n, low + 1, blocked_valid=False)
    assert np.array_equal(mo, em[om]) and np.array_equal(bo, eb[ob])


def test_n27_low14_windows_agree(gpu_ctx):
    """The config-5 shape: W = 28, LOW14, 4 moduli interleaved.  Groups 3..10
    swept as one window, then as two, give the same residues; seeded inputs
    are deterministic."""
    n, low, mods = 27, 14, list(PRIMES[:4])
    info, mc, bc = ofp.group_plan(n, 4, 3, 8, low)
    main = _inputs(info["main_count"], mods, 27)
    print("case n=27 groups 3..10 one window", flush=True)
    i1, whole, wblk = gpu_ctx.group_sweep(n, mods, 3, 8, low, main_in=main, blocked_out=True)
    assert i1["cells"] == 13 and i1["kernel_s"] > 0
    m0 = int(mc[:4].sum())
    b0 = int(bc[:4].sum())
    print("case n=27 groups 3..6 and 7..10", flush=True)
    _, a, ab = gpu_ctx.group_sweep(n, mods, 3, 4, low, main_in=main[:m0], blocked_out=True)
    _, b, bb = gpu_ctx.group_sweep(n, mods, 7, 4, low, main_in=main[m0:], blocked_out=True)
    assert np.array_equal(whole, np.concatenate([a, b])) and np.array_equal(wblk, np.concatenate([ab, bb]))
    assert not np.array_equal(whole, main)
    print("case n=27 seeded window", flush=True)
    s1, _, sb1 = gpu_ctx.group_sweep(n, mods, 100, 2, low, seed=5, blocked_out=True)
    s2, _, sb2 = gpu_ctx.group_sweep(n, mods, 100, 2, low, seed=5, blocked_out=True)
    assert np.array_equal(sb1, sb2) and s1["blocked_count"] == int(ofp.group_plan(n, 4, 100, 2, low)[2].sum())
