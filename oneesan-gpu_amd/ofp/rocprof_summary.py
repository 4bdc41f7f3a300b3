"""Summarise rocprofv3 CSV output for the cell kernels.

  python -m ofp.rocprof_summary trace   <kernel_trace.csv>          [--cells C]
  python -m ofp.rocprof_summary counter <counter_collection.csv>    [--cells C] [--scale F]

`trace` prints per-kernel call counts and mean durations.  `counter` prints the
mean counter value per kernel (KiB per dispatch for FETCH_SIZE / WRITE_SIZE)
and, with --scale, the calibrated bytes (the scale is what a device-to-device
copy of known size reads for the same counter).

With --cells C (cells per row, W-1) the cell launches of each moduli count KC
(full-state `cell_kernel` and height-capped `cell_capped_kernel` together, in
dispatch order, one launch per cell when KC <= 16) are cut into rows of C
launches: `trace` lists each row's time and how many of its cells ran capped,
then the per-cell times (p = W-1 .. 1) of the median full-state row; `counter`
gives the mean over full-state rows' dispatches and the per-cell values of the
last full-state row.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re
import statistics


def _short(name: str) -> str:
    name = name.replace("ofp::(anonymous namespace)::", "")
    return name.split("(ofp::")[0].split("(unsigned")[0]


def _kc(name: str) -> int | None:
    m = re.search(r"cell_(?:capped_)?kernel<(\d+)", name)
    return int(m.group(1)) if m else None


def _rows(seq, cells):
    """seq: [(order, value, capped)] -> rows of `cells` consecutive dispatches."""
    seq = sorted(seq)
    return [seq[i:i + cells] for i in range(0, len(seq) - cells + 1, cells)]


def trace(path: str, cells: int | None) -> None:
    rows = list(csv.DictReader(open(path)))
    by = collections.defaultdict(list)
    fam = collections.defaultdict(list)
    for r in rows:
        name = _short(r["Kernel_Name"])
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        by[name].append((t1 - t0) / 1e6)
        kc = _kc(name)
        if kc is not None:
            fam[kc].append((t0, (t1 - t0) / 1e6, "capped" in name))
    print(f"{'kernel':48s} {'calls':>6s} {'mean_ms':>10s} {'min_ms':>9s} {'max_ms':>9s}")
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k[:48]:48s} {len(v):6d} {sum(v) / len(v):10.3f} {min(v):9.3f} {max(v):9.3f}")
    if not cells:
        return
    for kc, seq in sorted(fam.items(), reverse=True):
        rws = _rows(seq, cells)
        if not rws:
            continue
        print(f"KC={kc}: {len(rws)} rows of {cells} cells; row ms (capped cells):")
        print("  " + " ".join(f"{sum(x[1] for x in r):.1f}({sum(x[2] for x in r)})" for r in rws))
        full = [r for r in rws if not any(x[2] for x in r)]
        if full:
            full.sort(key=lambda r: sum(x[1] for x in r))
            med = full[len(full) // 2]
            print(f"  median full-state row {sum(x[1] for x in med):.1f} ms, per-cell ms p={cells}..1: "
                  + " ".join(f"{x[1]:.1f}" for x in med))
            print(f"  mean full-state launch {statistics.fmean(x[1] for r in full for x in r):.3f} ms "
                  f"over {sum(len(r) for r in full)} launches")


def counter(path: str, scale: float, cells: int | None) -> None:
    rows = list(csv.DictReader(open(path)))
    agg = collections.defaultdict(list)
    fam = collections.defaultdict(list)
    name = None
    for i, r in enumerate(rows):
        kname = _short(r["Kernel_Name"])
        val = float(r["Counter_Value"])
        agg[kname].append(val)
        name = r["Counter_Name"]
        kc = _kc(kname)
        if kc is not None:
            order = int(r["Dispatch_Id"]) if r.get("Dispatch_Id") else i
            fam[kc].append((order, val, "capped" in kname))
    print(f"counter {name} (KiB per dispatch), calibration scale {scale}")
    for k, v in sorted(agg.items(), key=lambda kv: -sum(kv[1])):
        mean = sum(v) / len(v)
        print(f"{k[:48]:48s} dispatches={len(v):5d} mean_KiB={mean:14.1f} calibrated_GB={mean * 1024 * scale / 1e9:10.3f}")
    if not cells:
        return
    for kc, seq in sorted(fam.items(), reverse=True):
        full = [r for r in _rows(seq, cells) if not any(x[2] for x in r)]
        if not full:
            continue
        vals = [x[1] for r in full for x in r]
        mean = statistics.fmean(vals)
        print(f"KC={kc}: full-state rows {len(full)}, mean per dispatch {mean:.1f} KiB = "
              f"{mean * 1024 * scale / 1e9:.3f} GB calibrated")
        print(f"  last full-state row, calibrated GB per cell p={cells}..1: "
              + " ".join(f"{x[1] * 1024 * scale / 1e9:.1f}" for x in full[-1]))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter"])
    ap.add_argument("csv")
    ap.add_argument("--cells", type=int, default=None, help="cells per row (W-1) for the per-row listing")
    ap.add_argument("--scale", type=float, default=1.0)
    a = ap.parse_args(argv)
    if a.mode == "trace":
        trace(a.csv, a.cells)
    else:
        counter(a.csv, a.scale, a.cells)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
