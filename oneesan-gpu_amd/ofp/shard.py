"""One GPU's share of an n=27 row at real shape (measurement).

    python -m ofp.shard [--n 27] [--k 4] [--slices 64] [--slice 21]   one low-cells window
    python -m ofp.shard --share 0 [--shares 8]                        a whole row of one share

The n=27 frontier (520.7e9 states, 2 TB per modulus) does not fit one GPU; an
8-GPU run gives each GPU 1/8 of it (242 GiB per modulus, ~1 TB at 4 moduli),
which does not fit either, so a share is streamed through HBM in windows.  A
row has two halves, each over its own placement of the state (the reference's
bucket layout, docs/b300-bucket-snake.md:7-20):

  * high cells p = 27..15 (W = 28) over LOW14 occupancy groups (words with the
    same occupancy of positions 0..13; the high cells keep it, so a group is
    closed), 16,384 groups;
  * low cells p = 14..1 over HIGH13 occupancy groups (positions 15..27), 8,192
    groups, reading the blocked words the high half leaves.

A window is a run of consecutive groups, each stored alone with 32-bit ranks
local to the group, so every read and write of a window stays in it.  A share
is 1/`shares` of each half's groups by word count (consecutive masks).

`measure_share` streams every window of one share through both halves,
seeding each window in HBM, and reports the kernel time of the share's whole
row.  Between the halves, each LOW window's words are written into staging
buckets, one per share of the HIGH groups (ofp_group_exchange): the records
for the other shares are the send side of the per-row all-to-all.  `measure`
times one window of the older prefix-class placement of the low cells (cells
14..1 over 1/`slices` of the state, 64-bit global ranks, main words of other
shards read as seeded residues).

Left out (stated in the report): sending the staging buckets to the other GPUs
(the all-to-all itself; one GPU here), and the host or peer traffic that would
stage the non-resident windows (each window here is seeded in place, and the
low half's windows are seeded rather than assembled from the buckets).

Roofline bytes per launch: 8 B x k x (main + blocked words of the window).
"""
from __future__ import annotations

import json
import sys
import time

from . import _lib
from .path_bound import PRIMES

HBM_PEAK_GBS = 8000.0


def measure(ctx: _lib.Context, n: int = 27, k: int = 4, slices: int = 64, slice_: int | None = None,
            seed: int = 20261202) -> dict:
    if slice_ is None:
        slice_ = slices // 3
    mods = PRIMES[:k]
    info, _, _ = ctx.shard_sweep(n, mods, slices, slice_, seed=seed)
    states = info["main_count"] + info["blocked_count"]
    cells = info["low_cells"]
    launches = info["launches"]
    avg = info["kernel_s"] / launches
    declared = 8.0 * k * states
    nm, nb = _lib.state_counts(n)
    share = (nm + nb) / 8
    return {
        "workload": f"n={n} low cells {cells}..1, slice {slice_} of {slices} (whole prefix classes), "
                    f"{k} moduli interleaved, seeded inputs",
        "n": n, "k": k, "slices": slices, "slice": slice_, "low_cells": cells,
        "window_main_states": info["main_count"], "window_blocked_states": info["blocked_count"],
        "window_bytes": info["state_bytes"],
        "kernel_s": info["kernel_s"], "launches": launches, "avg_launch_ms": avg * 1e3,
        "transitions_per_s": cells * states * k / info["kernel_s"],
        "roofline": {"bound": "hbm", "achieved": declared / avg / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": declared / avg / 1e9 / HBM_PEAK_GBS, "bytes_per_launch": declared},
        "share_low_half_s_extrapolated": info["kernel_s"] * share / states,
        "stray_stores": info["stray_stores"],
        "left_out": "high cells 27..15 (transposed placement) and the per-row all-to-all exchange",
    }


def share_windows(n: int, k: int, shares: int, share: int, window_bytes: float, low: int = 0,
                  high: bool = False):
    """The windows of one share of a half (LOW groups / high cells, or HIGH
    groups / low cells): (first_group, groups) runs of consecutive occupancy
    groups whose device bytes stay <= window_bytes (a single larger group gets
    a window of its own).  The share holds the groups from the first whose
    words start at or after share/shares of the state."""
    import numpy as np
    W = n + 1
    low = low or W // 2
    fixed = W - 1 - low if high else low
    info, mc, bc = _lib.group_plan(n, k, 0, 1 << fixed, low, high)
    words = mc.astype(np.float64) + bc.astype(np.float64)
    cum = np.concatenate([[0.0], np.cumsum(words)])
    total = cum[-1]
    starts = [int(np.searchsorted(cum, total * s / shares, side="left")) for s in range(shares)] + [1 << fixed]
    g0, g1 = starts[share], starts[share + 1]
    per_word = 4 * (k + (k & 1) if k > 1 else 1)
    table = 15360  # one packed 32-bit group table
    windows, start, acc = [], g0, 0.0
    for g in range(g0, g1):
        b = words[g] * per_word + table
        if g > start and acc + b > window_bytes:
            windows.append((start, g - start))
            start, acc = g, 0.0
        acc += b
    if g1 > start:
        windows.append((start, g1 - start))
    return low, windows


def measure_share(ctx: _lib.Context, n: int = 27, k: int = 4, shares: int = 8, share: int = 0,
                  window_bytes: float = 136e9, seed: int = 20261202, log=None, max_windows: int = 0,
                  exchange: bool = True, exchange_window_bytes: float = 104e9) -> dict:
    """Stream every window of one share through both halves of a row
    (max_windows > 0: only the first windows of each half, for profiling runs).
    exchange: after the high half of each LOW window, write its words into the
    staging buckets of the shares owning their HIGH groups (ofp_group_exchange);
    the staging sits beside the window in HBM, so those windows are cut to
    exchange_window_bytes."""
    mods = PRIMES[:k]
    W = n + 1
    t0 = time.perf_counter()
    halves = {}
    xch = {"windows": 0, "words": 0, "kernel_s": 0.0, "moved_bytes": 0, "own_records": 0, "cross_records": 0,
           "max_staging_bytes": 0}
    for name, high in (("high", False), ("low", True)):
        wb = exchange_window_bytes if (exchange and not high) else window_bytes
        low, windows = share_windows(n, k, shares, share, wb, high=high)
        if max_windows > 0:
            windows = windows[:max_windows]
        cells = low if high else W - 1 - low
        h = {"placement": (f"HIGH{W - 1 - low} groups (occupancy of positions {low + 1}..{W - 1})" if high
                           else f"LOW{low} groups (occupancy of positions 0..{low - 1})"),
             "cells": f"{low}..1" if high else f"{W - 1}..{low + 1}", "windows": len(windows), "groups": 0,
             "words": 0, "kernel_s": 0.0, "launches": 0, "steps": 0, "max_window_bytes": 0,
             "cell_ms": [0.0] * cells}
        for g0, g in windows:
            info, _, _ = ctx.group_sweep(n, mods, g0, g, low, seed=seed, high=high)
            if exchange and not high:
                x, _ = ctx.group_exchange(n, mods, g0, g, low, shares=shares)
                xch["windows"] += 1
                xch["words"] += x["main_count"] + x["blocked_count"]
                xch["kernel_s"] += x["kernel_s"]
                xch["moved_bytes"] += x["moved_bytes"]
                own = x["main_records"][share] + x["blocked_records"][share]
                xch["own_records"] += own
                xch["cross_records"] += x["main_count"] + x["blocked_count"] - own
                xch["max_staging_bytes"] = max(xch["max_staging_bytes"], x["staging_bytes"])
                if log:
                    log(f"share {share}: exchange of groups {g0}..{g0 + g - 1}: {x['kernel_s']:.3f} s, "
                        f"{x['moved_bytes'] / x['kernel_s'] / 1e9:.0f} GB/s")
            h["groups"] += g
            h["words"] += info["main_count"] + info["blocked_count"]
            h["kernel_s"] += info["kernel_s"]
            h["launches"] += info["launches"]
            h["steps"] = info["steps"]  # passes over each window (a pair pass counts once)
            h["max_window_bytes"] = max(h["max_window_bytes"], info["state_bytes"])
            h["cell_ms"] = [a + b for a, b in zip(h["cell_ms"], info["cell_ms"])]
            if log:
                log(f"share {share}: {name} cells, groups {g0}..{g0 + g - 1}: {info['state_bytes'] / 1e9:.1f} GB, "
                    f"{info['kernel_s']:.3f} s")
        # declared bytes: one read and one write of every residue per pass over
        # the window (a pair pass, which applies two cells, counts once)
        declared = 8.0 * k * h["words"] * h["steps"]
        h["declared_bytes"] = declared
        h["transitions_per_s"] = cells * h["words"] * k / h["kernel_s"]
        h["roofline"] = {"bound": "hbm", "achieved": declared / h["kernel_s"] / 1e9, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": declared / h["kernel_s"] / 1e9 / HBM_PEAK_GBS}
        halves[name] = h
    wall = time.perf_counter() - t0
    nm, nb = _lib.state_counts(n)
    hi, lo = halves["high"], halves["low"]
    if exchange and xch["kernel_s"] > 0:
        gbps = xch["moved_bytes"] / xch["kernel_s"] / 1e9
        xch["roofline"] = {"bound": "hbm", "achieved": gbps, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": gbps / HBM_PEAK_GBS,
                           "bytes": "per word: its residues read (4 B x stride) + its record written "
                                    "(8 B index + 4 B x stride)"}
        xch["what"] = (f"each LOW window's words after the high half into {shares} staging buckets by the share "
                       f"owning their HIGH group; the cross-share records stand for the all-to-all's send "
                       f"buffers, the own share's for its local placement")
    return {
        "workload": f"n={n} row of one GPU's share ({share} of {shares}), {k} moduli interleaved, seeded "
                    f"inputs: high cells {hi['cells']} over {hi['placement']} in {hi['windows']} windows, "
                    f"low cells {lo['cells']} over {lo['placement']} in {lo['windows']} windows",
        "n": n, "k": k, "shares": shares, "share": share,
        "share_fraction_high": hi["words"] / (nm + nb), "share_fraction_low": lo["words"] / (nm + nb),
        "row_kernel_s": hi["kernel_s"] + lo["kernel_s"],
        "row_wall_s": wall,
        "transitions_per_s": (hi["transitions_per_s"] * hi["kernel_s"] + lo["transitions_per_s"] * lo["kernel_s"])
        / (hi["kernel_s"] + lo["kernel_s"]),
        "high": hi, "low": lo,
        "exchange": xch if exchange else None,
        "left_out": ("the transfer of the staging buckets to the other GPUs (the all-to-all itself), and the "
                     "host/peer staging of the non-resident windows (each window is seeded in HBM); wall time "
                     "includes the seeding") if exchange else
                    ("the per-row all-to-all between the two placements, and the host/peer staging of the "
                     "non-resident windows (each window is seeded in HBM); wall time includes the seeding"),
    }


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=27)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--slice", type=int, default=None)
    ap.add_argument("--share", type=int, default=None, help="stream a whole row of this share")
    ap.add_argument("--shares", type=int, default=8)
    ap.add_argument("--max-windows", type=int, default=0, help="only the first windows of each half (profiling)")
    ap.add_argument("--no-exchange", action="store_true", help="leave out the exchange after the high half")
    ap.add_argument("--exchange", action="store_true",
                    help="with --window 0 ...: after the sweep, exchange the window's words into --shares buckets")
    ap.add_argument("--window", type=int, nargs=3, metavar=("HIGH", "FIRST", "GROUPS"), default=None,
                    help="one window only: half (0: high cells / LOW groups, 1: low cells / HIGH groups), "
                         "first group and group count (profiling)")
    a = ap.parse_args(argv)
    with _lib.Context() as ctx:
        if a.window is not None:
            high, g0, g = a.window
            info, _, _ = ctx.group_sweep(a.n, PRIMES[:a.k], g0, g, 0, seed=20261202, high=bool(high))
            if a.exchange and not high:
                info["exchange"], _ = ctx.group_exchange(a.n, PRIMES[:a.k], g0, g, 0, shares=a.shares)
            print(json.dumps(info), flush=True)
            return 0
        if a.share is not None:
            rep = measure_share(ctx, a.n, a.k, a.shares, a.share,
                                log=lambda m: print(m, file=sys.stderr, flush=True), max_windows=a.max_windows,
                                exchange=not a.no_exchange)
            stray = 0
        else:
            rep = measure(ctx, a.n, a.k, a.slices, a.slice)
            stray = rep["stray_stores"]
    print(json.dumps(rep), flush=True)
    return 0 if stray == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
