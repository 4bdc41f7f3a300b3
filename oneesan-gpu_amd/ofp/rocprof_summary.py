"""Summarise rocprofv3 CSV output for the cell kernel.

  python -m ofp.rocprof_summary trace  <kernel_trace.csv>
  python -m ofp.rocprof_summary counter <counter_collection.csv> [--scale F]

`trace` prints per-kernel call counts and mean durations, plus the per-cell
(p = W-1 .. 1) durations of the first full row of each cell-kernel variant.
`counter` prints the mean counter value per kernel (KiB per dispatch for
FETCH_SIZE / WRITE_SIZE) and, with --scale, the calibrated bytes (the scale is
what a device-to-device copy of known size reads for the same counter).
"""
from __future__ import annotations

import argparse
import collections
import csv


def _short(name: str) -> str:
    name = name.replace("ofp::(anonymous namespace)::", "")
    return name.split("(ofp::")[0].split("(unsigned")[0]


def trace(path: str, cells: int | None) -> None:
    rows = list(csv.DictReader(open(path)))
    by = collections.defaultdict(list)
    for r in rows:
        by[_short(r["Kernel_Name"])].append((int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e6)
    print(f"{'kernel':48s} {'calls':>6s} {'mean_ms':>10s} {'min_ms':>9s} {'max_ms':>9s}")
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k[:48]:48s} {len(v):6d} {sum(v) / len(v):10.3f} {min(v):9.3f} {max(v):9.3f}")
    if cells:
        for k, v in by.items():
            if "cell_kernel" in k and len(v) >= 2 * cells:
                row = v[cells:2 * cells]  # second row (first is graph capture warm)
                print(f"{k}: per-cell ms, p={cells}..1: " + " ".join(f"{x:.1f}" for x in row))


def counter(path: str, scale: float) -> None:
    rows = list(csv.DictReader(open(path)))
    agg = collections.defaultdict(list)
    name = None
    for r in rows:
        agg[_short(r["Kernel_Name"])].append(float(r["Counter_Value"]))
        name = r["Counter_Name"]
    print(f"counter {name} (KiB per dispatch), calibration scale {scale}")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        mean = sum(v) / len(v)
        print(f"{k[:48]:48s} dispatches={len(v):5d} mean_KiB={mean:14.1f} calibrated_GB={mean * 1024 * scale / 1e9:10.3f}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter"])
    ap.add_argument("csv")
    ap.add_argument("--cells", type=int, default=None, help="cells per row (W-1) for the per-cell listing")
    ap.add_argument("--scale", type=float, default=1.0)
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv, a.cells)
    else:
        counter(a.csv, a.scale)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
