"""Group sweeps on the GPU through the C ABI: the high cells W-1..low+1 of a
row over occupancy groups of the low positions (LOW groups) and the low cells
low..1 over occupancy groups of the high positions (HIGH groups), each group
stored alone with 32-bit ranks local to the group.  Checked bit-exact against
the oracle's full-state cells, against the full-state GPU path, and, at the
n = 27 LOW14 / HIGH13 shape, against the same groups swept in other windows."""
from __future__ import annotations

import numpy as np
import pytest

import ofp
from group_masks import group_masks, occupancy_masks, window_order
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


@pytest.mark.parametrize("n,low,k", [(9, 4, 3), (12, 9, 1), (12, 5, 3), (13, 2, 2), (13, 7, 8), (11, 6, 14)])
def test_high_group_sweep_equals_oracle(gpu_ctx, oracle, n, low, k):
    """The low cells low..1 over HIGH groups, with blocked input (the high
    half's output), against the oracle's full-state cells."""
    mods = ([2] + list(PRIMES))[:k] if k > 1 else [PRIMES[1]]
    nm, nb = ofp.state_counts(n)
    fixed = n - low
    mm, mb = group_masks(n, low, True)
    om, ob = window_order(mm, 0, 1 << fixed), window_order(mb, 0, 1 << fixed)
    main = _inputs(nm, mods, 2000 * n + low)
    blk = _inputs(nb, mods, 3000 * n + low)
    print(f"case high n={n} low={low} k={k} all groups", flush=True)
    info, mo, _ = gpu_ctx.group_sweep(n, mods, 0, 1 << fixed, low, main_in=main[om], high=True,
                                      blocked_in=blk[ob])
    assert info["cells"] == low and len(info["cell_ms"]) == low
    for j, m in enumerate(mods):
        em, eb = np.ascontiguousarray(main[:, j]), np.ascontiguousarray(blk[:, j])
        oracle.apply_cells(n, m, em, eb, low, 1, True, pull=True)
        assert np.array_equal(mo[:, j], em[om]), (j, m)


def test_row_of_group_halves_equals_full_state_gpu(gpu_ctx):
    """n = 16 split at 8, 2 moduli: a whole row as the high half over LOW
    groups, a transpose to HIGH-group order on the host, and the low half over
    HIGH groups equals the full-state GPU row (ofp_apply_cells, pair passes)."""
    n, low, mods = 16, 8, [PRIMES[0], PRIMES[5]]
    nm, nb = ofp.state_counts(n)
    main = _inputs(nm, mods, 161)
    lm, lb = group_masks(n, low, False)
    hm, hb = group_masks(n, low, True)
    olm, olb = window_order(lm, 0, 1 << low), window_order(lb, 0, 1 << low)
    ohm, ohb = window_order(hm, 0, 1 << (n - low)), window_order(hb, 0, 1 << (n - low))
    print("case n=16 high half over LOW groups", flush=True)
    _, m1, b1 = gpu_ctx.group_sweep(n, mods, 0, 1 << low, low, main_in=main[olm], blocked_out=True)
    gm, gb = np.zeros_like(main), np.zeros((nb, 2), np.uint32)
    gm[olm], gb[olb] = m1, b1
    print("case n=16 low half over HIGH groups", flush=True)
    _, m2, _ = gpu_ctx.group_sweep(n, mods, 0, 1 << (n - low), low, main_in=gm[ohm], high=True,
                                   blocked_in=gb[ohb])
    got = np.zeros_like(main)
    got[ohm] = m2
    em, eb = main.copy(), np.zeros((nb, 2), np.uint32)
    print("case n=16 full-state row", flush=True)
    gpu_ctx.apply_cells(n, mods, em, eb, n, 1, blocked_valid=False)
    assert np.array_equal(got, em)


def test_n27_high13_windows_agree(gpu_ctx):
    """W = 28, HIGH13, 4 moduli: groups 0..7 as one window and as two give the
    same residues, with host inputs; seeded windows are deterministic."""
    n, low, mods = 27, 14, list(PRIMES[:4])
    info, mc, bc = ofp.group_plan(n, 4, 0, 8, low, high=True)
    main = _inputs(info["main_count"], mods, 270)
    blk = _inputs(info["blocked_count"], mods, 271)
    print("case n=27 HIGH groups 0..7 one window", flush=True)
    i1, whole, _ = gpu_ctx.group_sweep(n, mods, 0, 8, low, main_in=main, high=True, blocked_in=blk)
    assert i1["cells"] == 14 and i1["kernel_s"] > 0
    m0, b0 = int(mc[:3].sum()), int(bc[:3].sum())
    print("case n=27 HIGH groups 0..2 and 3..7", flush=True)
    _, a, _ = gpu_ctx.group_sweep(n, mods, 0, 3, low, main_in=main[:m0], high=True, blocked_in=blk[:b0])
    _, b, _ = gpu_ctx.group_sweep(n, mods, 3, 5, low, main_in=main[m0:], high=True, blocked_in=blk[b0:])
    assert np.array_equal(whole, np.concatenate([a, b])) and not np.array_equal(whole, main)
    print("case n=27 HIGH seeded window", flush=True)
    s1, _, _ = gpu_ctx.group_sweep(n, mods, 4000, 2, low, seed=9, high=True)
    assert s1["launches"] > 0 and len(s1["cell_ms"]) == 14


@pytest.mark.parametrize("high", [False, True])
def test_n27_pair_passes_equal_single_cells(high):
    """W = 28, 4 moduli: a LOW14 (or HIGH13) window swept as pair passes with each
    group's class tables equals the same window swept one cell per launch, with
    host inputs (the blocked input of the low half too)."""
    n, low, mods = 27, 14, list(PRIMES[:4])
    g0, g = (3, 4) if not high else (5, 3)
    info, _, _ = ofp.group_plan(n, 4, g0, g, low, high=high)
    main = _inputs(info["main_count"], mods, 2700 + high)
    blk = _inputs(info["blocked_count"], mods, 2800 + high) if high else None
    out = {}
    for pairs in (True, False):
        with ofp.Context() as ctx:
            ctx.set_group_launch(0, 0xFFFFFFFF if pairs else 0)
            print(f"case n=27 high={high} pairs={pairs}", flush=True)
            i, mo, bo = ctx.group_sweep(n, mods, g0, g, low, main_in=main, high=high, blocked_in=blk,
                                        blocked_out=not high)
            out[pairs] = (i, mo, bo)
    (ip, mp, bp), (is_, ms, bs) = out[True], out[False]
    assert np.array_equal(mp, ms) and not np.array_equal(mp, main)
    if not high:
        assert np.array_equal(bp, bs)
    assert ip["launches"] != is_["launches"]
    assert len(ip["cell_ms"]) == len(is_["cell_ms"]) == (low if high else n - low)


@pytest.mark.parametrize("n,low,k,high", [(11, 5, 3, False), (12, 6, 2, False), (13, 7, 4, True),
                                           (12, 5, 1, True), (11, 5, 14, True)])
def test_group_pair_passes_equal_oracle(oracle, n, low, k, high):
    """Group sweeps with pair passes (each group's class tables) against the
    oracle's full-state cells, both halves."""
    mods = ([2] + list(PRIMES))[:k] if k > 1 else [PRIMES[3]]
    nm, nb = ofp.state_counts(n)
    W = n + 1
    fixed = n - low if high else low
    mm, mb = group_masks(n, low, high)
    om, ob = window_order(mm, 0, 1 << fixed), window_order(mb, 0, 1 << fixed)
    main = _inputs(nm, mods, 4000 * n + low)
    blk = _inputs(nb, mods, 5000 * n + low) if high else None
    with ofp.Context() as ctx:
        ctx.set_group_launch(0, 0xFFFFFFFF)
        print(f"case pairs n={n} low={low} k={k} high={high}", flush=True)
        info, mo, bo = ctx.group_sweep(n, mods, 0, 1 << fixed, low, main_in=main[om], high=high,
                                       blocked_in=blk[ob] if high else None, blocked_out=not high)
    assert len(info["cell_ms"]) == (low if high else n - low)
    for j, m in enumerate(mods):
        em = np.ascontiguousarray(main[:, j])
        eb = np.ascontiguousarray(blk[:, j]) if high else np.zeros(nb, np.uint32)
        if high:
            oracle.apply_cells(n, m, em, eb, low, 1, True, pull=True)
        else:
            oracle.apply_cells(n, m, em, eb, W - 1, low + 1, False, pull=True)
        assert np.array_equal(mo[:, j], em[om]), (j, m)
        if not high:
            assert np.array_equal(bo[:, j], eb[ob]), (j, m)


@pytest.mark.parametrize("high", [False, True])
def test_group_streams_same_results(high):
    """A window of many small and a few large groups (n = 27 shape) swept with its
    group launches on one, two and four streams (pair passes, single cells,
    pair passes): bit-identical outputs."""
    n, low, mods = 27, 14, list(PRIMES[:4])
    first, groups = (2484, 12) if not high else (97, 12)
    ginfo, _, _ = ofp.group_plan(n, 4, first, groups, low, high=high)
    main = _inputs(ginfo["main_count"], mods, 4242 + high)
    outs = []
    with ofp.Context() as ctx:
        for ns in (1, 2, 4):
            ctx.set_group_launch(ns, 0 if ns == 2 else -1)
            print(f"case n=27 {'HIGH' if high else 'LOW'} groups {first}+{groups} streams={ns}", flush=True)
            _, mo, bo = ctx.group_sweep(n, mods, first, groups, low, main_in=main, blocked_out=not high, high=high)
            outs.append((mo, bo))
    assert not np.array_equal(outs[0][0], main)  # the sweep changed the window
    for mo, bo in outs[1:]:
        assert np.array_equal(mo, outs[0][0])
        if not high:
            assert np.array_equal(bo, outs[0][1])
