"""One GPU's share of an n=27 row at real shape (measurement).

    python -m ofp.shard [--n 27] [--k 4] [--slices 64] [--slice 21]

The n=27 frontier (520.7e9 states, 2 TB per modulus) does not fit one GPU; an
8-GPU run gives each GPU 1/8 of it (242 GiB per modulus, ~1 TB at 4 moduli),
which does not fit either.  This measures the owner-local part of a row at the
real shape: the low cells p = 14..1 (W = 28) over one resident window of
1/`slices` of the state (whole prefix classes, ~130 GB at slices=64, k=4),
with the real n=27 index math (64-bit ranks), 4 moduli interleaved, window
residues seeded uniform in [0, p), and main words of other shards (read only by
RR closures that flip a position above the low cells) read as seeded synthetic
residues.  The 1/8 share is 8 such windows; its time is extrapolated as 8x.

Left out (stated in the report): the high cells p = 27..15, which need the
transposed (LOW-occupancy-major) placement, and the per-row all-to-all that
moves the state between the two placements.

Roofline bytes per launch: 8 B x k x (main + blocked words of the window).
"""
from __future__ import annotations

import json
import sys

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


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=27)
    ap.add_argument("--k", type=int, default=4)
    ap.add_argument("--slices", type=int, default=64)
    ap.add_argument("--slice", type=int, default=None)
    a = ap.parse_args(argv)
    with _lib.Context() as ctx:
        rep = measure(ctx, a.n, a.k, a.slices, a.slice)
    print(json.dumps(rep), flush=True)
    return 0 if rep["stray_stores"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
