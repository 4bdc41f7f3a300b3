"""Launch-shape autotuner for the cell sweep.

For a given (n, k) it prepares the device state once, then times row sweeps
(cells W-1..1) for each candidate (threads per workgroup, workgroups per CU)
and keeps the fastest.  Timing runs advance the state, which does not change
the cost of a row as long as height caps are off (they are, for these stages).
A last stage picks the height-cap threshold on the first rows of a solve, the
ones caps shorten.  Results persist in a JSON cache keyed by device name, n
and k, so a solver or the serving loop can look them up without re-tuning.

Reference analogue: scripts/run/autotune.py:28 (the reference tunes its
frontier thread count and batch limits by re-running the solver binary);
GRIDFP_FRONTIER_THREADS / GRIDFP_FRONTIER_BLOCKS (row6crt20.cu:1182,1192,1306-1307).
"""
from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Iterable

from . import _lib

DEFAULT_CACHE = Path(os.environ.get("OFP_TUNE_CACHE", Path.home() / ".cache" / "ofp" / "tune.json"))
# Table shipped with the package (written by this module on an MI355X and
# committed); consulted when the user cache has no entry.
PACKAGED = Path(__file__).resolve().parent / "tuned.json"
THREADS = (128, 256)
BLOCKS_PER_CU = (4, 8, 16, 32)
LAYOUTS = (0, 1)
# second stage: cells p <= low_cells advance at most low_kc moduli per launch
LOW_CELLS = ((0, 16), (6, 8), (4, 8))
VARIANTS = (0, 1, 2, 3)
# chunks are dealt to the 8 XCDs in pieces of 2**shift chunks (-1: one eighth each)
XCD_SHIFTS = (-1, 0, 5, 8)
# full-state cells p >= this run the sorted kernel (0: none)
SORTED_FROM = (0, 4, 6, 8, 12)
# third stage: cap a cell when it then sweeps at most this % of the blocked words;
# timed on rows 1..CAPPED_ROWS of a solve (every row the forward bound caps at n <= 22)
CAP_PERCENTS = (75, 85, 95)
CAPPED_ROWS = 12


def _key(n: int, k: int) -> str:
    return f"n={n},k={k}"


def load_cache(path: Path = DEFAULT_CACHE) -> dict:
    try:
        return json.loads(path.read_text())
    except (OSError, ValueError):
        return {}


def save_cache(data: dict, path: Path = DEFAULT_CACHE) -> None:
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_suffix(".tmp")
    tmp.write_text(json.dumps(data, indent=1, sort_keys=True))
    os.replace(tmp, path)


def lookup(n: int, k: int, path: Path = DEFAULT_CACHE) -> dict | None:
    """Best known launch settings for (n, k): user cache first, then the packaged
    table (exact key, else the entry of the same kind -- one modulus or several --
    nearest in n, then in k)."""
    ent = load_cache(path).get(_key(n, k))
    if ent:
        return ent
    pk = load_cache(PACKAGED)
    if _key(n, k) in pk:
        return pk[_key(n, k)]
    cands = []
    for key, v in pk.items():
        kn, kk = int(key.split(",")[0][2:]), int(key.split(",")[1][2:])
        if (kk > 1) == (k > 1):
            cands.append(((abs(kn - n), abs(kk - k), kn, kk), v))
    return min(cands, key=lambda t: t[0])[1] if cands else None


def pair_mask_for(n: int, k: int, path: Path = DEFAULT_CACHE) -> int:
    """The pair mask to run (n, k) with: the one tuned for (n, k), else for the
    same n and the nearest moduli count (one modulus or several first), else 0
    (the library's default).  A mask names cell positions, which shift with n, so
    one tuned at another n does not carry over."""
    key = _key(n, k)
    tables = (load_cache(path), load_cache(PACKAGED))
    for table in tables:
        if key in table:
            return int(table[key].get("pair_mask", 0))
    for table in tables:
        same_n = [(((kk > 1) != (k > 1)), abs(kk - k), kk, v) for kk, v in
                  ((int(kk_.split(",")[1][2:]), v_) for kk_, v_ in table.items()
                   if int(kk_.split(",")[0][2:]) == n)]
        if same_n:
            return int(min(same_n, key=lambda t: t[:3])[3].get("pair_mask", 0))
    return 0


def apply(ctx: "_lib.Context", n: int, k: int) -> dict | None:
    """Set the best known launch shape and layout on ctx (before prepare/solve)."""
    ent = lookup(n, k)
    if ent:
        ctx.set_layout(int(ent.get("layout", 1)))
        ctx.set_launch(int(ent["threads"]), int(ent["blocks_per_cu"]))
        ctx.set_variant(int(ent.get("variant", 2)))
        ctx.set_low_cells(int(ent.get("low_cells", 0)), int(ent.get("low_kc", 16)))
        ctx.set_height_caps(int(ent.get("cap_percent", _lib.DEFAULT_CAP_PERCENT)))
        ctx.set_xcd_pieces(int(ent.get("xcd_shift", 5)))
        ctx.set_sorted_cells(int(ent.get("sorted_from", -1)))
        ctx.set_pair_passes(int(ent.get("pair_mode", 1)), pair_mask_for(n, k))
    return ent


def pair_stage(ctx: "_lib.Context", n: int, moduli: list[int], log=None) -> tuple[int, float]:
    """Which cells pair up: one row with one launch per cell and one with a pair
    pass over every pair (p, p-1), p = W-1, W-3, ..., both timed per launch on
    whole-state rows; a pair is kept where its pass beats its two cells.  Returns
    (cell mask, the row's time in ms with that choice); mask 0: no pair pays."""
    ctx.set_height_caps(0)
    ctx.set_pair_passes(0)
    ctx.prepare(n, moduli)
    ctx.time_cells(n)
    single = {p: ms for p, ms, _ in ctx.time_cells(n)}
    ctx.set_pair_passes(1, 0xFFFFFFFF)
    ctx.time_cells(n)
    paired = {p: ms for p, ms, name in ctx.time_cells(n) if name == "pair_kernel"}
    mask, row = 0, 0.0
    W = n + 1
    p = W - 1
    while p >= 1:
        if p in paired and p - 1 >= 1 and paired[p] < single[p] + single[p - 1]:
            mask |= 1 << p
            row += paired[p]
            p -= 2
        else:
            row += single[p]
            p -= 1
    if log:
        log(f"autotune n={n} k={len(moduli)} pairs: mask {mask:#x}, row {row:.2f} ms "
            f"(one launch per cell {sum(single.values()):.2f} ms)")
    return mask, row


def tune(ctx: "_lib.Context", n: int, moduli: list[int], rows: int = 1,
         threads: Iterable[int] = THREADS, blocks_per_cu: Iterable[int] = BLOCKS_PER_CU,
         layouts: Iterable[int] = LAYOUTS, path: Path | None = DEFAULT_CACHE, log=None) -> dict:
    """Time every (layout, launch shape); set the best on ctx; return the table."""
    table = []
    best = None
    ctx.set_height_caps(0)
    ctx.set_variant(2)
    ctx.set_low_cells(0, 16)
    ctx.set_xcd_pieces(5)
    ctx.set_sorted_cells(0)
    ctx.set_pair_passes(0)  # the launch-shape stages time the per-cell kernels
    for lay in layouts:
        ctx.set_layout(lay)
        ctx.prepare(n, moduli)
        for t in threads:
            for b in blocks_per_cu:
                ctx.set_launch(t, b)
                ctx.row_sweep(1)  # capture the row graph for this shape
                t0 = time.perf_counter()
                st = ctx.row_sweep(rows)
                el = time.perf_counter() - t0
                row_ms = el / rows * 1e3
                ent = {"layout": lay, "threads": t, "blocks_per_cu": b, "row_ms": row_ms,
                       "kernel_ms_per_row": st["kernel_s"] / rows * 1e3}
                table.append(ent)
                if log:
                    log(f"autotune n={n} k={len(moduli)} layout={lay} threads={t} blocks/cu={b}: "
                        f"{row_ms:.2f} ms/row")
                if best is None or row_ms < best["row_ms"]:
                    best = ent
    ctx.set_layout(best["layout"])
    ctx.set_launch(best["threads"], best["blocks_per_cu"])
    best = dict(best, low_cells=0, low_kc=16, variant=2)
    if best["layout"] == 1:
        ctx.prepare(n, moduli)
        for var in VARIANTS:
            for lc, lk in (LOW_CELLS if len(moduli) > 1 else ((0, 16),)):
                ctx.set_variant(var)
                ctx.set_low_cells(lc, lk)
                ctx.row_sweep(1)
                t0 = time.perf_counter()
                st = ctx.row_sweep(rows)
                row_ms = (time.perf_counter() - t0) / rows * 1e3
                ent = dict(best, variant=var, low_cells=lc, low_kc=lk, row_ms=row_ms,
                           kernel_ms_per_row=st["kernel_s"] / rows * 1e3)
                table.append(ent)
                if log:
                    log(f"autotune n={n} k={len(moduli)} variant={var} low_cells={lc} low_kc={lk}: "
                        f"{row_ms:.2f} ms/row")
                if row_ms < best["row_ms"]:
                    best = ent
    ctx.set_variant(best["variant"])
    ctx.set_low_cells(best["low_cells"], best["low_kc"])
    best = dict(best, xcd_shift=5)
    ctx.prepare(n, moduli)
    for sh in XCD_SHIFTS:
        ctx.set_xcd_pieces(sh)
        ctx.row_sweep(1)
        t0 = time.perf_counter()
        st = ctx.row_sweep(rows)
        row_ms = (time.perf_counter() - t0) / rows * 1e3
        ent = dict(best, xcd_shift=sh, row_ms=row_ms, kernel_ms_per_row=st["kernel_s"] / rows * 1e3)
        table.append(ent)
        if log:
            log(f"autotune n={n} k={len(moduli)} xcd_shift={sh}: {row_ms:.2f} ms/row")
        if row_ms < best["row_ms"]:
            best = ent
    ctx.set_xcd_pieces(best["xcd_shift"])
    best = dict(best, sorted_from=0)
    if best["layout"] == 1:
        ctx.prepare(n, moduli)
        for sf in SORTED_FROM:
            ctx.set_sorted_cells(sf)
            ctx.row_sweep(1)
            t0 = time.perf_counter()
            st = ctx.row_sweep(rows)
            row_ms = (time.perf_counter() - t0) / rows * 1e3
            ent = dict(best, sorted_from=sf, row_ms=row_ms, kernel_ms_per_row=st["kernel_s"] / rows * 1e3)
            table.append(ent)
            if log:
                log(f"autotune n={n} k={len(moduli)} sorted_from={sf}: {row_ms:.2f} ms/row")
            if sf == 0 or row_ms < best["row_ms"]:  # sf = 0 re-times the baseline on this state
                best = ent
    ctx.set_sorted_cells(best["sorted_from"])
    best = dict(best, pair_mode=0, pair_mask=0)
    if best["layout"] == 1 and n <= 22:
        mask, row = pair_stage(ctx, n, moduli, log)
        table.append(dict(best, pair_mode=1 if mask else 0, pair_mask=mask, row_ms=row))
        if mask:
            best = dict(best, pair_mode=1, pair_mask=mask, row_ms=row)
    ctx.set_pair_passes(best["pair_mode"], best["pair_mask"])
    best = dict(best, cap_percent=_lib.DEFAULT_CAP_PERCENT)
    rows_capped = min(CAPPED_ROWS, n - 1)
    if best["layout"] == 1 and rows_capped >= 1:
        timed = {}
        for pct in CAP_PERCENTS:
            ctx.set_height_caps(pct)
            ctx.prepare(n, moduli)
            ctx.row_sweep(1)
            t0 = time.perf_counter()
            ctx.row_sweep(rows_capped)
            timed[pct] = (time.perf_counter() - t0) / rows_capped * 1e3
            table.append(dict(best, cap_percent=pct, capped_rows_ms=timed[pct]))
            if log:
                log(f"autotune n={n} k={len(moduli)} cap_percent={pct}: rows 1..{rows_capped} "
                    f"{timed[pct]:.2f} ms/row")
        best["cap_percent"] = min(timed, key=timed.get)
    ctx.set_height_caps(best["cap_percent"])
    result = {"best": best, "table": table}
    if path is not None:
        data = load_cache(path)
        data[_key(n, len(moduli))] = {"layout": best["layout"], "threads": best["threads"],
                                      "blocks_per_cu": best["blocks_per_cu"],
                                      "low_cells": best["low_cells"], "low_kc": best["low_kc"],
                                      "variant": best["variant"],
                                      "cap_percent": best["cap_percent"],
                                      "xcd_shift": best["xcd_shift"],
                                      "sorted_from": best["sorted_from"],
                                      "pair_mode": best["pair_mode"], "pair_mask": best["pair_mask"],
                                      "row_ms": best["row_ms"], "tuned_at": time.time()}
        save_cache(data, path)
    return result


def main(argv=None) -> int:
    import argparse
    import sys
    from .path_bound import PRIMES
    ap = argparse.ArgumentParser(description="tune the cell-sweep launch shape")
    ap.add_argument("--n", type=int, nargs="+", default=[22])
    ap.add_argument("--k", type=int, nargs="+", default=[1, 14])
    ap.add_argument("--rows", type=int, default=1)
    ap.add_argument("--json", type=str, default="")
    ap.add_argument("--cache", type=str, default=str(DEFAULT_CACHE))
    args = ap.parse_args(argv)
    out = {}
    with _lib.Context() as ctx:
        for n in args.n:
            for k in args.k:
                res = tune(ctx, n, PRIMES[:k], rows=args.rows, path=Path(args.cache),
                           log=lambda m: print(m, file=sys.stderr, flush=True))
                out[_key(n, k)] = res
                print(json.dumps({"n": n, "k": k, "best": res["best"]}), flush=True)
    if args.json:
        Path(args.json).write_text(json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
