"""Summarise rocprofv3 CSV output for the cell kernels.

  python -m ofp.rocprof_summary trace   <kernel_trace.csv>          [--cells C]
  python -m ofp.rocprof_summary counter <counter_collection.csv>    [--cells C] [--scale F]
  python -m ofp.rocprof_summary valu    <counter_collection.csv>    --cells C [--cus N] [--xcds X]
  python -m ofp.rocprof_summary dram    <counter_collection.csv>

`dram` reads a run collected with
  --pmc TCC_EA0_RDREQ_DRAM_32B_sum TCC_EA0_WRREQ_WRITE_DRAM_32B_sum
(read and write requests to DRAM in 32-byte units, a 64-byte request counted
twice and a 128-byte one four times) and prints, per cell-kernel family, the
mean DRAM bytes read and written per dispatch.

`trace` prints per-kernel call counts and mean durations; it reads the kernel
trace as CSV (--output-format csv) or as the SQLite database rocprofv3 writes by
default (*_results.db).  `counter` prints the
mean counter value per kernel (KiB per dispatch for FETCH_SIZE / WRITE_SIZE)
and, with --scale, the calibrated bytes (the scale is what a device-to-device
copy of known size reads for the same counter).

With --cells C (cells per row, W-1) the launches of each moduli count KC are
cut into rows of C cells (--rows R keeps the first R), in dispatch order, one
launch per cell when KC <= 16: a full-state `cell_kernel` / `cell_sorted_kernel`
or height-capped `cell_capped_kernel` launch is one cell, and the consecutive
`pair_kernel` launches of one pair pass (one per height change d, in the order
-1, 0, 1, -2, 2; a pass that has no classes of some d skips it) are two cells.
`trace` lists each row's time and how many of its cells ran capped, then the
per-step times (p = W-1 .. 1; a pair pass as "p,p-1") of the median full-state
row; `counter` gives the mean over full-state rows' dispatches and the per-cell
values of the last full-state row (single-cell launches only).

`valu` reads one run collected with
  --pmc SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_VALU SQ_INSTS_SALU
        SQ_INSTS_LDS SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_WAVES GRBM_GUI_ACTIVE GRBM_COUNT
and prints, per cell of the last full-state row, the integer-VALU figures:
  busy  = SQ_ACTIVE_INST_VALU / CUs / (GRBM_GUI_ACTIVE / XCDs): the share of the
          CUs' VALU issue cycles used (one wave64 VALU instruction per CU per cycle;
          GRBM_GUI_ACTIVE is summed over the XCDs);
  lanes = SQ_THREAD_CYCLES_VALU / (SQ_ACTIVE_INST_VALU * 64): active lanes per
          VALU instruction (divergence);
and VALU / SALU / LDS / VMEM instruction counts per 64 blocked words
(SQ_WAVES counts the persistent waves, so per-word figures use --words).
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
    m = re.search(r"(?:cell_(?:capped_|sorted_)?|pair_)kernel<(\d+)", name)
    return int(m.group(1)) if m else None


_PAIR_ORDER = {-1: 0, 0: 1, 1: 2, -2: 3, 2: 4, -3: 5, 3: 6}


def _pair_d(name: str) -> int | None:
    """Height change d of a pair_kernel launch (its third template argument)."""
    m = re.search(r"pair_kernel<\d+,\s*[^,]+,\s*(-?\d+)", name)
    return int(m.group(1)) if m else None


def _steps(seq):
    """seq: [(t0, t1, capped, d)] (ns; d None for a one-cell launch) -> steps
    [(t0, ms, capped, ncells)]: a one-cell launch is a step of one cell, the
    pair launches of one pass a step of two.  A pass's launches may run side by
    side (its readers d = -1, 0, 1 on three streams, then d = -2, 2), so a step's
    time is its wall span, first start to last end, and a new pass starts where a
    launch's d is already in the current one (every d occurs once per pass)."""
    out = []
    seen = None  # d's of the current pass
    for t0, t1, capped, d in sorted(seq):
        if d is None:
            out.append([t0, t1, capped, 1])
            seen = None
            continue
        if seen is not None and d not in seen:
            out[-1][1] = max(out[-1][1], t1)
            out[-1][2] = out[-1][2] or capped
            seen.add(d)
        else:
            out.append([t0, t1, capped, 2])
            seen = {d}
    return [(a, (b - a) / 1e6, c, n) for a, b, c, n in out]


def _step_rows(seq, cells):
    """Steps cut into rows of `cells` cells."""
    rows, cur, n = [], [], 0
    for st in _steps(seq):
        cur.append(st)
        n += st[3]
        if n >= cells:
            rows.append(cur)
            cur, n = [], 0
    return rows[:MAX_ROWS] if MAX_ROWS else rows


def kernel_family(name: str) -> str | None:
    """'cell_sorted_kernel<14>', 'pair_kernel<14>' etc. for a cell or pair kernel's
    full name (index width, window change and other template arguments dropped),
    None for any other kernel."""
    m = re.search(r"((?:cell_(?:capped_|sorted_|shard_)?|pair_)kernel)<(\d+)", name)
    return f"{m.group(1)}<{m.group(2)}>" if m else None


MAX_ROWS = None  # only the first rows of each moduli count (--rows); None = all


def _rows(seq, cells):
    """seq: [(order, value, capped)] -> rows of `cells` consecutive dispatches."""
    seq = sorted(seq)
    rows = [seq[i:i + cells] for i in range(0, len(seq) - cells + 1, cells)]
    return rows[:MAX_ROWS] if MAX_ROWS else rows


def _trace_rows(path: str) -> list[dict]:
    """Kernel dispatches as {Kernel_Name, Start_Timestamp, End_Timestamp} (ns)."""
    if path.endswith(".db"):
        import sqlite3
        con = sqlite3.connect(path)
        try:
            got = con.execute("select name, start, end from kernels order by start").fetchall()
        finally:
            con.close()
        return [{"Kernel_Name": n, "Start_Timestamp": s, "End_Timestamp": e} for n, s, e in got]
    return list(csv.DictReader(open(path)))


def trace(path: str, cells: int | None, kcs: set[int] | None = None) -> None:
    rows = _trace_rows(path)
    by = collections.defaultdict(list)
    fam = collections.defaultdict(list)
    for r in rows:
        name = _short(r["Kernel_Name"])
        t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        by[name].append((t1 - t0) / 1e6)
        kc = _kc(name)
        if kc is not None:
            fam[kc].append((t0, t1, "capped" in name, _pair_d(name)))
    print(f"{'kernel':48s} {'calls':>6s} {'mean_ms':>10s} {'min_ms':>9s} {'max_ms':>9s}")
    for k, v in sorted(by.items(), key=lambda kv: -sum(kv[1])):
        print(f"{k[:48]:48s} {len(v):6d} {sum(v) / len(v):10.3f} {min(v):9.3f} {max(v):9.3f}")
    if not cells:
        return
    for kc, seq in sorted(fam.items(), reverse=True):
        if kcs and kc not in kcs:
            continue
        rws = _step_rows(seq, cells)
        if not rws:
            continue
        print(f"KC={kc}: {len(rws)} rows of {cells} cells; row ms (capped cells):")
        print("  " + " ".join(f"{sum(x[1] for x in r):.1f}({sum(x[2] for x in r)})" for r in rws))
        full = [r for r in rws if not any(x[2] for x in r)]
        if full:
            full.sort(key=lambda r: sum(x[1] for x in r))
            med = full[len(full) // 2]
            parts, p = [], cells
            for x in med:
                parts.append(f"{p},{p - 1}:{x[1]:.1f}" if x[3] == 2 else f"{p}:{x[1]:.1f}")
                p -= x[3]
            print(f"  median full-state row {sum(x[1] for x in med):.1f} ms over {len(full)} full-state rows; "
                  f"per step (p: ms): " + " ".join(parts))


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


VALU_COUNTERS = ("SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU", "SQ_INSTS_VALU", "SQ_INSTS_SALU",
                 "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_WAVES", "GRBM_GUI_ACTIVE")


def valu(path: str, cells: int, cus: int, xcds: int, words: int | None) -> None:
    """Per step (a one-cell launch, or the launches of one pair pass, whose
    counters are summed) of the last full-state row of each moduli count."""
    by = collections.defaultdict(dict)
    meta = {}
    for r in csv.DictReader(open(path)):
        d = int(r["Dispatch_Id"])
        by[d][r["Counter_Name"]] = float(r["Counter_Value"])
        meta[d] = (_short(r["Kernel_Name"]), int(r["Start_Timestamp"]), int(r["End_Timestamp"]))
    fam = collections.defaultdict(list)
    for d, (name, t0, t1) in sorted(meta.items()):
        kc = _kc(name)
        if kc is not None and all(c in by[d] for c in VALU_COUNTERS):
            fam[kc].append((d, "capped" in name, _pair_d(name)))
    for kc, seq in sorted(fam.items(), reverse=True):
        # steps in dispatch order: [dispatches, capped, cells]
        steps, seen = [], None
        for d, capped, pd in seq:
            if pd is None:
                steps.append([[d], capped, 1])
                seen = None
            elif seen is not None and pd not in seen:
                steps[-1][0].append(d)
                seen.add(pd)
            else:
                steps.append([[d], capped, 2])
                seen = {pd}
        rows, cur, n = [], [], 0
        for st in steps:
            cur.append(st)
            n += st[2]
            if n >= cells:
                rows.append(cur)
                cur, n = [], 0
        if MAX_ROWS:
            rows = rows[:MAX_ROWS]
        full = [r for r in rows if not any(x[1] for x in r)]
        if not full:
            continue
        row = full[-1]
        print(f"KC={kc}: last full-state row ({len(full)} full-state rows); per step, top of the row first")
        print(f"{'p':>5s} {'ms':>7s} {'busy%':>6s} {'lanes%':>6s} {'VALU/64w':>9s} {'SALU/64w':>9s} "
              f"{'LDS/64w':>8s} {'VMRD/64w':>9s} {'VMWR/64w':>9s}")
        tot_valu = tot_cyc = 0.0
        p = cells
        for ds, _, nc in row:
            agg = collections.Counter()
            ms = 0.0
            for d in ds:
                for c in VALU_COUNTERS:
                    agg[c] += by[d][c]
                ms += (meta[d][2] - meta[d][1]) / 1e6
            cyc = agg["GRBM_GUI_ACTIVE"] / xcds
            busy = 100.0 * agg["SQ_ACTIVE_INST_VALU"] / cus / cyc
            lanes = 100.0 * agg["SQ_THREAD_CYCLES_VALU"] / (agg["SQ_ACTIVE_INST_VALU"] * 64)
            w64 = (words / 64.0) if words else agg["SQ_WAVES"]
            tot_valu += agg["SQ_ACTIVE_INST_VALU"]
            tot_cyc += cyc
            label = f"{p},{p - 1}" if nc == 2 else f"{p}"
            print(f"{label:>5s} {ms:7.2f} {busy:6.1f} {lanes:6.1f} "
                  f"{agg['SQ_INSTS_VALU'] / w64:9.0f} {agg['SQ_INSTS_SALU'] / w64:9.0f} {agg['SQ_INSTS_LDS'] / w64:8.1f} "
                  f"{agg['SQ_INSTS_VMEM_RD'] / w64:9.1f} {agg['SQ_INSTS_VMEM_WR'] / w64:9.1f}")
            p -= nc
        print(f"  row: VALU busy {100.0 * tot_valu / cus / tot_cyc:.1f}% of the CUs' VALU issue cycles")


DRAM_COUNTERS = ("TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")


def dram_bytes(path: str) -> dict[str, dict]:
    """{family: {"read": B, "write": B, "dispatches": d, "read_total": B,
    "write_total": B}}: means per dispatch and totals (a pair pass is seven
    dispatches, one per window height change)."""
    per: dict[int, dict] = {}
    for r in csv.DictReader(open(path)):
        e = per.setdefault(int(r["Dispatch_Id"]), {"kernel": r["Kernel_Name"]})
        e[r["Counter_Name"]] = float(r["Counter_Value"])
    out: dict[str, dict] = {}
    for d in sorted(per):
        e = per[d]
        fam = kernel_family(e["kernel"])
        if fam is None or not all(c in e for c in DRAM_COUNTERS):
            continue
        o = out.setdefault(fam, {"read": 0.0, "write": 0.0, "dispatches": 0})
        o["read"] += 32.0 * e[DRAM_COUNTERS[0]]
        o["write"] += 32.0 * e[DRAM_COUNTERS[1]]
        o["dispatches"] += 1
    for o in out.values():
        o["read_total"], o["write_total"] = o["read"], o["write"]
        o["read"] /= o["dispatches"]
        o["write"] /= o["dispatches"]
    return out


def dram(path: str) -> None:
    print(f"{'kernel':28s} {'dispatches':>10s} {'read_GB':>10s} {'write_GB':>10s} {'total_GB':>10s}")
    for fam, o in sorted(dram_bytes(path).items()):
        print(f"{fam:28s} {o['dispatches']:10d} {o['read'] / 1e9:10.3f} {o['write'] / 1e9:10.3f} "
              f"{(o['read'] + o['write']) / 1e9:10.3f}")


_PAIR_D = re.compile(r"pair_kernel<(\d+), [^,]+, (-?\d+),")
_ORDER = (-1, 0, 1, -2, 2)


def launches(path: str) -> None:
    """Every dispatch of a counter run with its DRAM bytes and duration; the
    pair-pass dispatches are grouped into passes (their launches run in the
    order d = -1, 0, 1, -2, 2, each present only where the pass has classes of
    that d; a new pass starts where the order restarts)."""
    per: dict[int, dict] = {}
    for r in csv.DictReader(open(path)):
        e = per.setdefault(int(r["Dispatch_Id"]), {"kernel": r["Kernel_Name"],
                                                   "ns": int(r["End_Timestamp"]) - int(r["Start_Timestamp"])})
        e[r["Counter_Name"]] = float(r["Counter_Value"])
    print(f"{'dispatch':>8s} {'kernel':22s} {'d':>3s} {'ms':>8s} {'read_GB':>8s} {'write_GB':>8s} {'TB/s':>6s}")
    pas = None
    prev_kc, prev_i = None, 99

    def close(p):
        if p and p["n"] > 1:
            print(f"{'':8s} {'  pass total':22s} {'':3s} {p['ms']:8.2f} {p['rd'] / 1e9:8.1f} {p['wr'] / 1e9:8.1f} "
                  f"{(p['rd'] + p['wr']) / p['ms'] / 1e9:6.2f}")
    for d in sorted(per):
        e = per[d]
        if not all(c in e for c in DRAM_COUNTERS):
            continue
        rd, wr, ms = 32.0 * e[DRAM_COUNTERS[0]], 32.0 * e[DRAM_COUNTERS[1]], e["ns"] / 1e6
        fam = kernel_family(e["kernel"])
        m = _PAIR_D.search(e["kernel"])
        if m:
            kc, dd = int(m.group(1)), int(m.group(2))
            i = _ORDER.index(dd) if dd in _ORDER else 99
            if kc != prev_kc or i <= prev_i:
                close(pas)
                pas = {"ms": 0.0, "rd": 0.0, "wr": 0.0, "n": 0}
            prev_kc, prev_i = kc, i
            pas["ms"] += ms
            pas["rd"] += rd
            pas["wr"] += wr
            pas["n"] += 1
            dtxt = str(dd)
        else:
            close(pas)
            pas, prev_kc, prev_i = None, None, 99
            dtxt = ""
        if fam is None or ms <= 0:
            continue
        print(f"{d:8d} {fam:22s} {dtxt:>3s} {ms:8.2f} {rd / 1e9:8.1f} {wr / 1e9:8.1f} {(rd + wr) / ms / 1e9:6.2f}")
    close(pas)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("mode", choices=["trace", "counter", "valu", "dram", "launches"])
    ap.add_argument("csv")
    ap.add_argument("--cells", type=int, default=None, help="cells per row (W-1) for the per-row listing")
    ap.add_argument("--rows", type=int, default=None,
                    help="only the first ROWS rows of each moduli count (one solve's W rows; later "
                         "launches of other sizes are left out)")
    ap.add_argument("--kc", type=str, default=None,
                    help="moduli per launch whose rows to list, comma separated (trace; default all)")
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--cus", type=int, default=256, help="compute units (MI355X: 256)")
    ap.add_argument("--xcds", type=int, default=8, help="XCDs GRBM_GUI_ACTIVE is summed over (MI355X: 8)")
    ap.add_argument("--words", type=int, default=None, help="blocked words per launch (per-64-word counts)")
    a = ap.parse_args(argv)
    global MAX_ROWS
    MAX_ROWS = a.rows
    if a.mode == "trace":
        trace(a.csv, a.cells, {int(x) for x in a.kc.split(",")} if a.kc else None)
    elif a.mode == "counter":
        counter(a.csv, a.scale, a.cells)
    elif a.mode == "dram":
        dram(a.csv)
    elif a.mode == "launches":
        launches(a.csv)
    else:
        if not a.cells:
            ap.error("valu needs --cells")
        valu(a.csv, a.cells, a.cus, a.xcds, a.words)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
