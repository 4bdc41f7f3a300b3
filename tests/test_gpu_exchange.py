"""The exchange between the two placements of a row on the GPU, through the C
ABI: a whole row of an n-grid run as it would run on `shares` GPUs -- every
share sweeps the high half over its LOW window, writes its words into staging
buckets by the share that owns their HIGH group (ofp_group_exchange), the
buckets are delivered, and every share sweeps the low half over its HIGH
window -- must equal the full-state row (ofp_apply_cells), bit-exact.  The
staging itself is checked against the high half's output permuted on the host.
At the n = 27 shape, LOW windows are exchanged and their buckets checked for
counts, ranges, the residues they carry, window order inside each bucket and
run-to-run identity."""
from __future__ import annotations

import numpy as np
import pytest

import ofp
from group_masks import group_masks, window_order
from ofp.path_bound import PRIMES

pytestmark = pytest.mark.gpu


def _inputs(nm: int, mods, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)


def _share_starts(counts: np.ndarray, shares: int) -> list[int]:
    """Share boundaries over consecutive masks cut by word count (ofp.shard's split)."""
    cum = np.concatenate([[0.0], np.cumsum(counts.astype(np.float64))])
    return [int(np.searchsorted(cum, cum[-1] * s / shares, side="left")) for s in range(shares)] + [len(counts)]


@pytest.mark.parametrize("n,low,k,shares", [(11, 5, 3, 4), (13, 7, 2, 8), (14, 6, 1, 3), (16, 8, 4, 8)])
def test_row_over_shares_equals_full_state(gpu_ctx, n, low, k, shares):
    mods = [2, PRIMES[0], PRIMES[3], PRIMES[7]][:k] if k > 1 else [PRIMES[2]]
    nm, nb = ofp.state_counts(n)
    W = n + 1
    main = _inputs(nm, mods, 100 * n + low)
    lm, lb = group_masks(n, low, False)
    hm, hb = group_masks(n, low, True)
    nl, nh = 1 << low, 1 << (W - 1 - low)
    _, lmc, lbc = ofp.group_plan(n, k, 0, nl, low)
    lstart = _share_starts(lmc + lbc, shares)
    # the high half's output in global order, share by share, and the staging buckets
    after_m, after_b = np.zeros_like(main), np.zeros((nb, k), np.uint32)
    recs = []
    hfirst = None
    for s in range(shares):
        g0, g = lstart[s], lstart[s + 1] - lstart[s]
        if g == 0:
            continue
        om, ob = window_order(lm, g0, g), window_order(lb, g0, g)
        print(f"case n={n} low={low} k={k} LOW share {s}: groups {g0}+{g}", flush=True)
        _, mo, bo = gpu_ctx.group_sweep(n, mods, g0, g, low, main_in=main[om], blocked_out=True)
        after_m[om], after_b[ob] = mo, bo
        info, b = gpu_ctx.group_exchange(n, mods, g0, g, low, shares=shares, download=True)
        assert info["main_count"] == len(om) and info["blocked_count"] == len(ob)
        assert sum(info["main_records"]) == len(om) and sum(info["blocked_records"]) == len(ob)
        assert info["kernel_s"] > 0
        if hfirst is None:
            hfirst = info["share_first"]
        assert info["share_first"] == hfirst
        recs.append((info, b))
    _, hmc, hbc = ofp.group_plan(n, k, 0, nh, low, high=True)
    assert hfirst == _share_starts(hmc + hbc, shares)
    # deliver: every share's HIGH placement from all buckets
    got = np.zeros_like(main)
    for s in range(shares):
        h0, h = hfirst[s], hfirst[s + 1] - hfirst[s]
        if h == 0:
            continue
        ohm, ohb = window_order(hm, h0, h), window_order(hb, h0, h)
        pm = np.zeros((len(ohm), k), np.uint32)
        pb = np.zeros((len(ohb), k), np.uint32)
        seen_m = np.zeros(len(ohm), np.int64)
        seen_b = np.zeros(len(ohb), np.int64)
        for info, b in recs:
            for kind, place, seen in (("main", pm, seen_m), ("blocked", pb, seen_b)):
                cnt = info[f"{kind}_records"]
                lo = sum(cnt[:s])
                idx = b[f"{kind}_index"][lo:lo + cnt[s]].astype(np.int64)
                assert idx.size == 0 or (idx.min() >= 0 and idx.max() < len(place))
                place[idx] = b[f"{kind}_res"][lo:lo + cnt[s]]
                np.add.at(seen, idx, 1)
        assert (seen_m == 1).all() and (seen_b == 1).all()
        assert np.array_equal(pm, after_m[ohm]) and np.array_equal(pb, after_b[ohb])
        print(f"case n={n} HIGH share {s}: groups {h0}+{h}", flush=True)
        _, mo, _ = gpu_ctx.group_sweep(n, mods, h0, h, low, main_in=pm, high=True, blocked_in=pb)
        got[ohm] = mo
    em, eb = main.copy(), np.zeros((nb, k), np.uint32)
    print(f"case n={n} full-state row", flush=True)
    gpu_ctx.apply_cells(n, mods, em, eb, n, 1, blocked_valid=False)
    assert np.array_equal(got, em)


def test_n27_window_buckets(gpu_ctx):
    """W = 28, LOW14 groups 3..6 with 4 moduli into 8 shares: bucket counts
    add up, indices stay inside their share's placement and are distinct, and
    the records carry the window's residues."""
    n, low, mods, shares = 27, 14, list(PRIMES[:4]), 8
    ginfo, mc, bc = ofp.group_plan(n, 4, 3, 4, low)
    main = _inputs(ginfo["main_count"], mods, 2727)
    print("case n=27 LOW groups 3..6 sweep", flush=True)
    _, mo, bo = gpu_ctx.group_sweep(n, mods, 3, 4, low, main_in=main, blocked_out=True)
    print("case n=27 LOW groups 3..6 exchange", flush=True)
    info, b = gpu_ctx.group_exchange(n, mods, 3, 4, low, shares=shares, download=True)
    assert sum(info["main_records"]) == len(mo) and sum(info["blocked_records"]) == len(bo)
    _, hmc, hbc = ofp.group_plan(n, 4, 0, 1 << 13, low, high=True)
    first = info["share_first"]
    assert first == _share_starts(hmc + hbc, shares)
    for kind, counts, src in (("main", hmc, mo), ("blocked", hbc, bo)):
        rec = info[f"{kind}_records"]
        idx = b[f"{kind}_index"]
        res = b[f"{kind}_res"]
        lo = 0
        for s in range(shares):
            part = idx[lo:lo + rec[s]]
            size = int(counts[first[s]:first[s + 1]].sum())
            assert part.size == 0 or int(part.max()) < size
            assert np.unique(part).size == part.size
            lo += rec[s]
        # the records carry exactly the window's residues (as a multiset per modulus)
        for j in range(4):
            assert np.array_equal(np.sort(res[:, j]), np.sort(src[:, j]))


def _keys(res: np.ndarray) -> np.ndarray:
    """A 64-bit key per row: its first two residues."""
    return (res[:, 0].astype(np.uint64) << np.uint64(32)) | res[:, 1].astype(np.uint64)


def test_n27_large_window_order_and_identity(gpu_ctx):
    """LOW14 groups 1000..1001 (25.4 M words, tiles of many rounds and blocks of
    every length): two exchanges of the same window give identical buckets; each
    record's residues are those of one word of the window (each word once), and
    inside a bucket the records follow the window order; indices are distinct and
    inside their share's placement."""
    n, low, mods, shares = 27, 14, list(PRIMES[:4]), 8
    ginfo, _, _ = ofp.group_plan(n, 4, 1000, 2, low)
    main = _inputs(ginfo["main_count"], mods, 2728)
    print("case n=27 LOW groups 1000..1001 sweep", flush=True)
    _, mo, bo = gpu_ctx.group_sweep(n, mods, 1000, 2, low, main_in=main, blocked_out=True)
    print("case n=27 LOW groups 1000..1001 exchange (twice)", flush=True)
    info, b = gpu_ctx.group_exchange(n, mods, 1000, 2, low, shares=shares, download=True)
    info2, b2 = gpu_ctx.group_exchange(n, mods, 1000, 2, low, shares=shares, download=True)
    assert info["main_records"] == info2["main_records"] and info["blocked_records"] == info2["blocked_records"]
    for key in b:
        assert np.array_equal(b[key], b2[key]), key
    _, hmc, hbc = ofp.group_plan(n, 4, 0, 1 << 13, low, high=True)
    first = info["share_first"]
    for kind, counts, src in (("main", hmc, mo), ("blocked", hbc, bo)):
        rec = info[f"{kind}_records"]
        idx, res = b[f"{kind}_index"], b[f"{kind}_res"]
        wk = _keys(src)
        order = np.argsort(wk)
        sk = wk[order]
        assert np.all(sk[1:] != sk[:-1])  # the seeded residues tell the words apart
        pos = np.searchsorted(sk, _keys(res))
        assert np.all(pos < sk.size)
        wpos = order[pos]
        assert np.array_equal(src[wpos], res)  # every record is a word of the window
        assert np.unique(wpos).size == wpos.size == src.shape[0]  # each word exactly once
        lo = 0
        for s in range(shares):
            part, wp = idx[lo:lo + rec[s]], wpos[lo:lo + rec[s]]
            size = int(counts[first[s]:first[s + 1]].sum())
            assert part.size == 0 or int(part.max()) < size
            assert np.unique(part).size == part.size
            assert np.all(np.diff(wp) > 0), (kind, s)  # window order inside the bucket
            lo += rec[s]


def test_exchange_needs_the_swept_window(gpu_ctx):
    n, low, mods = 12, 6, [PRIMES[0]]
    print("case exchange of a window not swept", flush=True)
    gpu_ctx.group_sweep(n, mods, 0, 4, low)
    with pytest.raises(ofp.OfpError):
        gpu_ctx.group_exchange(n, mods, 4, 4, low, shares=2)
    info, _ = gpu_ctx.group_exchange(n, mods, 0, 4, low, shares=2)
    assert sum(info["main_records"]) == info["main_count"]
