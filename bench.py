#!/usr/bin/env python3
"""Benchmark: frontier transitions/s and residues/hour, batched vs one-residue,
on one MI355X, for the corner-to-corner simple-path count at n=22.

Workload: n=22, k = the number of primes the checkerboard-strip bound needs
(14), the largest primes below 2^32 -- an exact-count (CRT) run.  Inputs: the
solve's own initial state (synthetic in the sense that nothing is read from
disk); every buffer is device-resident before any timed region starts.

Measured in one process, in this order:
  * solve: one complete n=22 solve with all k moduli interleaved (state
    initialisation + W = 23 rows; the memory is allocated before), timed whole.
    All k residues are reported; the first is checked against the reference's
    published value and every residue the committed ggcount fixtures hold for
    n=22 is checked too.  residues_per_hour = 3600 k / solve time.  The first
    rows and the last two rows of a solve sweep only the words of small height
    (height caps; the answer is unchanged), so the solve does less than W full
    rows of work.
  * steps (the `value`): --warmup untimed and --steps timed row sweeps of the
    same state with height caps off, so that every row does its full
    algorithmic work (W-1 cells x S states x k moduli transitions); each cell is
    launched on its own between HIP events on the library's stream, and the
    timed region is bracketed by device synchronisation.  value = transitions
    of the timed rows / their wall time.
  * roofline: per kernel (cell_sorted_kernel: cells p >= 6, cell_kernel: p <= 5)
    the mean launch duration over the timed rows' events; declared algorithmic
    bytes per launch = 8 B x k x S (one 4-byte read and one 4-byte write per
    state per modulus).  `roofline` is the dominant kernel's.  HBM traffic per
    launch is measured before the GPU work starts, by running one row of the same
    configuration in a child process under rocprofv3 with the DRAM request
    counters TCC_EA0_RDREQ_DRAM_32B and TCC_EA0_WRREQ_WRITE_DRAM_32B (32-byte
    units; no calibration needed); null if rocprofv3 is unavailable.  The
    height-capped kernel of the solve is reported beside them.
  * plain: the same solve with k=1 (one residue at a time), timed whole ->
    residues/hour of the plain form, and its per-kernel launch times.
  * plain_small: whole one-modulus solves at n=12 and n=18 (median of three).
  * shard_n27: one GPU's share of an n=27 row, both halves, every window of the
    share streamed through HBM, and between the halves the exchange of every
    LOW window's words into staging buckets by destination share
    (ofp.shard.measure_share).
  * cpu_baseline: the CPU oracle's pull-form step on a bounded sample of the
    n=22 cell step, its whole n=12 solve (median of three) and one whole n=16
    solve, on the host cores.

Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "oneesan-gpu_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
GOLDEN_FIRST = {22: (4294967291, 1245069298), 21: (4294967291, 998035516),
                20: (4294967291, 2308006916), 18: (4294967291, 503411004)}
COUNTERS = ("TCC_EA0_RDREQ_DRAM_32B_sum", "TCC_EA0_WRREQ_WRITE_DRAM_32B_sum")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model() -> str | None:
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def golden_residues(n: int) -> dict[int, int]:
    """Residues of n the committed fixtures hold: {modulus: residue}."""
    out: dict[int, int] = {}
    gd = ROOT / "tests" / "golden"
    for name in ("reference_residues.json", "ggcount_residues.json", "ggcount_residues_large.json"):
        try:
            rows = json.loads((gd / name).read_text())["residues"]
        except (OSError, ValueError, KeyError):
            continue
        for r in rows:
            if r["n"] == n:
                out[int(r["modulus"])] = int(r["residue"])
    return out


def capped_words(width: int, cap: int) -> int:
    """Words of `width` positions entered at height 1 whose heights stay <= cap."""
    d = [0] * (cap + 2)
    d[0] = 1  # words of width 0 end at height 0
    for _ in range(width):
        d = [(d[h] + (d[h - 1] if h > 0 else 0) + d[h + 1]) if h <= cap else 0 for h in range(cap + 1)] + [0]
    return d[1]


def cpu_baseline(n: int, seconds: float) -> dict:
    """Time the CPU oracle's pull-form cell step on a bounded sample of the
    n-sized state (consecutive target windows of a middle cell), at one
    modulus; rate = target states processed per second."""
    import ctypes

    from oracle import binding
    lib = binding.load()
    fn = lib.or_cell_pull_range
    fn.restype = None
    u32p = ctypes.POINTER(ctypes.c_uint32)
    fn.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, u32p, u32p, u32p, u32p,
                   ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]
    nm, nb = int(lib.or_main_count(n)), int(lib.or_blocked_count(n))
    libc = ctypes.CDLL("libc.so.6")
    libc.calloc.restype = ctypes.c_void_p
    libc.calloc.argtypes = [ctypes.c_size_t, ctypes.c_size_t]
    libc.free.argtypes = [ctypes.c_void_p]
    bufs = [libc.calloc(nm, 4), libc.calloc(nb, 4), libc.calloc(nm, 4), libc.calloc(nb, 4)]
    if not all(bufs):
        for b in bufs:
            if b:
                libc.free(b)
        return {"value": None, "unit": "transitions/s", "cores": 0, "kind": "port",
                "sample": "allocation failed"}
    mi, bi, mo, bo = (ctypes.cast(b, u32p) for b in bufs)
    W = n + 1
    p = W // 2
    win_m, win_b = 1 << 20, 1 << 19
    t0m, t0b = nm // 3, nb // 3
    done = 0
    t_start = time.perf_counter()
    windows = 0
    while time.perf_counter() - t_start < seconds and t0m + win_m <= nm and t0b + win_b <= nb:
        fn(n, p, 4294967291, mi, bi, mo, bo, t0m, t0m + win_m, t0b, t0b + win_b)
        done += win_m + win_b
        t0m += win_m
        t0b += win_b
        windows += 1
    el = time.perf_counter() - t_start
    for b in bufs:
        libc.free(b)
    cores = int(lib.or_num_threads())
    rate = done / el if el > 0 else None
    return {"value": rate, "unit": "transitions/s", "cores": cores, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": (f"oracle pull-form step (oracle/gridfp_oracle.c or_cell_pull_range), n={n} cell "
                       f"p={p}, {windows} windows of {win_m} main + {win_b} blocked target states from "
                       f"rank 1/3 on, one modulus, {el:.1f} s on {cores} threads; per-residue CPU "
                       f"time for k moduli = k x this"),
            "residues_per_hour_extrapolated": (3600.0 * rate / (W * (W - 1) * (nm + nb))) if rate else None}


def cpu_solve(n: int, mod: int, repeats: int, warmup: bool = True, progress: bool = False) -> dict:
    """The CPU oracle's whole solve (pull form, one modulus): one warm-up (if
    warmup) and the median of `repeats` when repeats > 1, else one timed solve;
    progress: a log line after each solve."""
    from oracle import binding
    if repeats > 1 and warmup:
        binding.solve(n, mod)
    times, res = [], None
    for i in range(repeats):
        t0 = time.perf_counter()
        res = binding.solve(n, mod)
        times.append(time.perf_counter() - t0)
        if progress:
            log(f"bench: cpu n={n} solve {i + 1}/{repeats}: {times[-1]:.1f} s")
    med = sorted(times)[len(times) // 2]
    W = n + 1
    lib = binding.load()
    S = int(lib.or_main_count(n)) + int(lib.or_blocked_count(n))
    return {"solve_s": med, "residue": res, "transitions_per_s": W * (W - 1) * S / med,
            "sample": f"oracle whole n={n} solve (oracle/gridfp_oracle.c or_solve, pull form), "
                      + (f"median of {repeats}" if repeats > 1 else "one run")}


def copy_rate(nbytes: int) -> dict | None:
    """One device-to-device copy of nbytes (the HIP runtime's copy kernel): the
    practical streaming ceiling next to the published 8 TB/s."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return None
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(src), ctypes.c_size_t(nbytes)) != 0:
        return None
    if hip.hipMalloc(ctypes.byref(dst), ctypes.c_size_t(nbytes)) != 0:
        hip.hipFree(src)
        return None
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)
    hip.hipDeviceSynchronize()
    best = None
    for _ in range(3):
        t0 = time.perf_counter()
        rc = hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
        hip.hipDeviceSynchronize()
        el = time.perf_counter() - t0
        if rc == 0 and (best is None or el < best):
            best = el
    hip.hipFree(src)
    hip.hipFree(dst)
    if best is None:
        return None
    return {"bytes": nbytes, "seconds": best, "GBps_read_plus_write": 2 * nbytes / best / 1e9,
            "what": "hipMemcpy device-to-device, best of 3 (host clock around each copy)"}


def counter_child(n: int, k: int) -> int:
    """Run under rocprofv3 by measure_traffic(): one whole-state row of the bench
    configuration (k moduli), then one of the plain form (k=1), each cell launched
    on its own."""
    import ofp
    from ofp import autotune
    from ofp.path_bound import PRIMES, primes_for_n
    mods = primes_for_n(n) if k == 0 else PRIMES[:k]
    with ofp.Context(device=0) as ctx:
        for kk in (len(mods), 1):
            autotune.apply(ctx, n, kk)
            ctx.set_height_caps(0)
            ctx.prepare(n, mods[:kk])
            ctx.time_cells(n)
    print(json.dumps({"counter_child": "done", "n": n, "k": len(mods)}), flush=True)
    return 0


def measure_traffic(n: int, k: int, timeout: float) -> dict:
    """HBM bytes per launch for each cell kernel, from the DRAM request counters of
    one row run in a child process under rocprofv3 (before this process touches the
    GPU).  Returns {"kernels": {name: {"read": B, "write": B, "dispatches": d}}} or
    {"error": ...}."""
    prof = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    if not os.path.exists(prof):
        return {"error": "rocprofv3 not found"}
    tmp = tempfile.mkdtemp(prefix="ofp_ctr_", dir="/tmp")
    env = dict(os.environ, TMPDIR="/tmp")
    cmd = [prof, "--pmc", *COUNTERS, "-d", tmp, "-o", "ctr", "--output-format", "csv", "--",
           sys.executable, str(Path(__file__).resolve()), "--counter-child", "--n", str(n), "--k", str(k)]
    log(f"bench: counters: {' '.join(cmd[:4])} ... (one row of n={n}, k={k} and k=1 under rocprofv3)")
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env, cwd="/tmp")
    except subprocess.TimeoutExpired:
        shutil.rmtree(tmp, ignore_errors=True)
        return {"error": f"rocprofv3 child exceeded {timeout:.0f} s"}
    files = glob.glob(os.path.join(tmp, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        shutil.rmtree(tmp, ignore_errors=True)
        return {"error": f"no counter output (rc={r.returncode}): {r.stderr[-400:]}"}
    from ofp.rocprof_summary import dram_bytes
    out: dict[str, dict] = {}
    for path in files:
        for fam, o in dram_bytes(path).items():
            out[fam] = o
    shutil.rmtree(tmp, ignore_errors=True)
    return {"kernels": out, "counters": list(COUNTERS), "unit": "bytes per dispatch (counter x 32 B)",
            "rc": r.returncode}


def kernel_rows(cell_times: list[list[tuple[int, float, str]]], kc: int, states: int, k: int,
                traffic: dict | None) -> dict:
    """Per-kernel statistics over rows of time_cells() output: one entry per cell
    launched on its own (one launch when k <= 16; a cell of k > 16 moduli sums its
    launches) and one per pair pass (cells p, p-1; its five launches, one per
    height change d = -2..2 over the two cells, together).  Traffic per pair pass = the pair kernel's
    DRAM bytes over one row / the passes in a row."""
    by: dict[str, list[float]] = {}
    cells: dict[str, set] = {}
    for row in cell_times:
        for p, ms, name in row:
            if name == "pair_second":
                continue  # the second cell of a pair pass (timed with the first)
            by.setdefault(name, []).append(ms)
            cells.setdefault(name, set()).add(p)
    passes_per_row = len(by.get("pair_kernel", [])) / max(1, len(cell_times))
    declared = 8.0 * states * k
    out = {}
    for name, v in by.items():
        avg = sum(v) / len(v)
        ent = {"kernel": f"{name}<{kc}>", "cells": sorted(cells[name], reverse=True), "launches": len(v),
               "avg_launch_ms": avg, "bytes_per_launch": declared, "achieved": declared / (avg * 1e-3) / 1e9,
               "unit": "GB/s"}
        ent["frac"] = ent["achieved"] / HBM_PEAK_GBS
        t = (traffic or {}).get(f"{name}<{kc}>")
        if name == "pair_kernel":
            ent["cells"] = [f"{p},{p - 1}" for p in ent["cells"]]
            ent["unit"] = "GB/s per pass (two cells, five launches)"
            if t and passes_per_row:
                t = {"read": t["read_total"] / passes_per_row, "write": t["write_total"] / passes_per_row}
        if t:
            tb = t["read"] + t["write"]
            ent.update({"traffic": tb, "traffic_read": t["read"], "traffic_write": t["write"],
                        "traffic_GBps": tb / (avg * 1e-3) / 1e9,
                        "traffic_frac": tb / (avg * 1e-3) / 1e9 / HBM_PEAK_GBS,
                        "traffic_over_declared": tb / declared})
        out[name] = ent
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed full-work row sweeps")
    ap.add_argument("--warmup", type=int, default=2, help="untimed row sweeps before timing")
    ap.add_argument("--n", type=int, default=22)
    ap.add_argument("--k", type=int, default=0, help="moduli (0 = as many as the CRT bound needs)")
    ap.add_argument("--no-plain", action="store_true", help="skip the k=1 plain-form solve")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-n18", action="store_true",
                    help="also time the CPU oracle's whole n=18 solve (median of three; a few minutes)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-shard", action="store_true", help="skip the n=27 one-GPU-share measurement")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the whole plain-form solves at n=12 and n=18 (k=1)")
    ap.add_argument("--no-counters", action="store_true", help="skip the rocprofv3 traffic measurement")
    ap.add_argument("--counter-timeout", type=float, default=240.0)
    ap.add_argument("--counter-child", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: the frontier state of one solve "
                                   "lives in one GPU's HBM; run with --gpus 1"}), flush=True)
        return 2
    if args.counter_child:
        return counter_child(args.n, args.k)

    import ofp
    from ofp.path_bound import PRIMES, primes_for_n

    n = args.n
    W = n + 1
    mods = primes_for_n(n) if args.k == 0 else PRIMES[:args.k]
    k = len(mods)
    nm, nb = ofp.state_counts(n)
    S = nm + nb
    log(f"bench: n={n} W={W} k={k} states={S} ({nm} main + {nb} blocked)")

    traffic = None
    if not args.no_counters:
        traffic = measure_traffic(n, k, args.counter_timeout)
        log(f"bench: counters {json.dumps(traffic)[:300]}")

    from ofp import autotune
    ctx = ofp.Context(device=0)
    tuned = autotune.apply(ctx, n, k)
    log(f"bench: launch config {tuned}")
    plan = ctx.plan(n, k)
    log(f"bench: state {plan.state_bytes / 2**30:.1f} GiB, free {plan.free_bytes / 2**30:.1f} GiB")

    # ---- one complete solve, all k moduli (the state memory is allocated first) ----
    ctx.prepare(n, mods)
    ts0 = time.perf_counter()
    ctx.prepare(n, mods)
    ss = ctx.row_sweep(W)
    residues = ctx.read_main(0, k)
    solve_s = time.perf_counter() - ts0
    golden = golden_residues(n)
    checked = {str(m): (r == golden[m]) for m, r in zip(mods, residues) if m in golden}
    verified = (residues[0] == GOLDEN_FIRST[n][1]) if n in GOLDEN_FIRST and GOLDEN_FIRST[n][0] == mods[0] else None
    residues_per_hour = 3600.0 * k / solve_s
    log(f"bench: whole solve {solve_s:.2f} s -> {residues_per_hour:.0f} residues/h; residues {residues}; "
        f"first verified={verified}; fixtures {checked}")
    capped = None
    if ss["capped_launches"]:
        # The solve's rows with height-capped cells, timed as runs of such rows (HIP
        # events around each run); their other cells run the full-state kernels.
        # Declared bytes: 8 B x k x (blocked words swept + main words within the cap)
        # for a capped cell, 8 B x k x S for a full-state one.
        pct = tuned.get("cap_percent", 85) if tuned else 85
        decl, ncapped = 0.0, 0
        for r in range(W):
            caps = [ofp.height_cap(n, r, p, pct) for p in range(W - 1, 0, -1)]
            if not any(c for c, _ in caps):
                continue
            for cap, nbc in caps:
                if cap:
                    decl += 8.0 * k * (nbc + capped_words(W, cap))
                    ncapped += 1
                else:
                    decl += 8.0 * k * S
        capped = {"kernel": f"rows with cell_capped_kernel<{min(k, 16)}> launches", "launches": ss["capped_launches"],
                  "capped_launches": ncapped * -(-k // 16),
                  "kernel_s": ss["capped_kernel_s"], "avg_launch_ms": ss["capped_kernel_s"] / ss["capped_launches"] * 1e3,
                  "bytes_total": decl, "achieved": decl / ss["capped_kernel_s"] / 1e9, "unit": "GB/s",
                  "frac": decl / ss["capped_kernel_s"] / 1e9 / HBM_PEAK_GBS,
                  "basis": "the solve's rows with capped cells (HIP events around runs of such rows); declared "
                           "bytes: capped cell 8 B x k x (blocked words swept + main words within the cap), "
                           "full-state cell 8 B x k x S"}

    # ---- timed steps: full-work rows (height caps off), each cell between events ----
    ctx.set_height_caps(0)
    ctx.prepare(n, mods)
    for _ in range(args.warmup):
        ctx.time_cells(n)
    log(f"bench: warmup {args.warmup} rows done")
    t0 = time.perf_counter()
    rows = [ctx.time_cells(n) for _ in range(args.steps)]
    el = time.perf_counter() - t0
    ctx.set_height_caps(True)
    kernel_ms = sum(ms for row in rows for _, ms, _ in row) * 1e-3
    transitions = args.steps * (W - 1) * S * k
    value = transitions / el
    log(f"bench: {args.steps} full rows in {el:.3f} s (kernel {kernel_ms:.3f} s)")
    kc = min(k, 16)
    kern = kernel_rows(rows, kc, S, k, (traffic or {}).get("kernels"))
    dom = max(kern.values(), key=lambda e: e["avg_launch_ms"] * e["launches"])
    roof = {"bound": "hbm", "achieved": dom["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom["frac"], "traffic": dom.get("traffic"), "kernel": dom["kernel"],
            "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_ms": dom["avg_launch_ms"],
            "launches": dom["launches"],
            "launch_basis": f"{dom['launches']} launches of the timed rows, HIP events on the library's stream",
            "traffic_source": ("rocprofv3 --pmc " + " ".join(COUNTERS) + " x 32 B, one row in a child process "
                               "of this run") if dom.get("traffic") else None}
    if dom.get("traffic"):
        roof["traffic_frac"] = dom["traffic_frac"]

    # ---- plain one-residue-at-a-time form (k=1): one whole solve, then one timed row ----
    plain = None
    if not args.no_plain:
        tuned1 = autotune.apply(ctx, n, 1)
        ctx.prepare(n, [mods[0]])
        tp0 = time.perf_counter()
        ctx.prepare(n, [mods[0]])
        sp = ctx.row_sweep(W)
        pres = ctx.read_main(0, 1)
        tp = time.perf_counter() - tp0
        ctx.set_height_caps(0)
        prow = [ctx.time_cells(n) for _ in range(2)]
        autotune.apply(ctx, n, 1)
        pk = kernel_rows(prow, 1, S, 1, (traffic or {}).get("kernels"))
        plain = {"solve_s": tp, "residues_per_hour": 3600.0 / tp, "transitions_per_s": sp["transitions"] / tp,
                 "residue": pres[0], "residue_matches_batched": pres[0] == residues[0],
                 "full_row_ms": sum(ms for _, ms, _ in prow[-1]), "kernels": pk, "launch": tuned1}
        log(f"bench: plain k=1 solve {tp:.2f} s -> {plain['residues_per_hour']:.0f} residues/h "
            f"(residue {pres[0]})")
    ctx.close()

    copy = copy_rate(16 << 30)

    small = None
    if not args.no_small:
        small = {}
        for ns in (12, 18):
            with ofp.Context(device=0) as c:
                autotune.apply(c, ns, 1)
                c.solve(ns, [PRIMES[0]])
                times, res = [], None
                for _ in range(3):
                    t0s = time.perf_counter()
                    r, _ = c.solve(ns, [PRIMES[0]])
                    times.append(time.perf_counter() - t0s)
                    res = r[0]
            med = sorted(times)[1]
            Ws = ns + 1
            Ss = sum(ofp.state_counts(ns))
            want = GOLDEN_FIRST.get(ns)
            small[f"n{ns}"] = {"solve_s": med, "residue": res, "modulus": PRIMES[0],
                               "verified": (res == want[1]) if want and want[0] == PRIMES[0] else None,
                               "transitions_per_s": Ws * (Ws - 1) * Ss / med,
                               "residues_per_hour": 3600.0 / med}
            log(f"bench: plain n={ns} solve {med * 1e3:.1f} ms, residue {res}")

    shard = None
    if not args.no_shard:
        from ofp import shard as shard_mod
        with ofp.Context(device=0) as sctx:  # the library's defaults (no sorted cells at 4 moduli)
            shard = shard_mod.measure_share(sctx, share=0)
        log(f"bench: n=27 share 0 of 8, whole row streamed in {shard['high']['windows']}+{shard['low']['windows']} "
            f"windows: {shard['row_kernel_s']:.2f} s kernel ({shard['row_wall_s']:.1f} s wall); high cells "
            f"{shard['high']['roofline']['frac']:.3f}, low cells {shard['low']['roofline']['frac']:.3f} of HBM peak")
        x = shard.get("exchange")
        if x:
            log(f"bench: n=27 share 0 exchange: {x['windows']} windows, {x['kernel_s']:.2f} s, "
                f"{x['roofline']['achieved']:.0f} GB/s ({x['roofline']['frac']:.3f} of peak), "
                f"{x['cross_records'] / max(x['words'], 1):.3f} of the records for other shares")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, args.cpu_seconds)
        log(f"bench: cpu baseline {cpu['value']:.3e} transitions/s on {cpu['cores']} threads")
        cpu["n12"] = cpu_solve(12, PRIMES[0], 3)
        cpu["n16"] = cpu_solve(16, PRIMES[0], 1)
        if args.cpu_n18:
            cpu["n18"] = cpu_solve(18, PRIMES[0], 3, warmup=False, progress=True)
            log(f"bench: cpu n=18 solve {cpu['n18']['solve_s']:.1f} s (median of 3), residue {cpu['n18']['residue']}")
        if small:
            cpu["n12"]["residue_matches_gpu"] = cpu["n12"]["residue"] == small["n12"]["residue"]
            if "n18" in cpu:
                cpu["n18"]["residue_matches_gpu"] = cpu["n18"]["residue"] == small["n18"]["residue"]

    out = {
        "metric": "frontier transitions/s & residues/hour, batched vs 1-residue, n=22, bit-exact",
        "value": value,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (the solve's own initial frontier; nothing loaded)",
        "config": {"workload": f"exact CRT count, n={n} ({W}x{W} vertices), {k} moduli interleaved; a step is "
                               f"one full-work row sweep (height caps off)", "n": n, "moduli": k, "states": S,
                   "launch": tuned,
                   "main_states": nm, "blocked_states": nb, "state_bytes": plan.state_bytes,
                   "parallelism": "single GPU"},
        "value_basis": "transitions of the timed full-work rows (W-1 cells x S states x k moduli each) / "
                       "their wall time",
        "residues_per_hour": residues_per_hour,
        "residues_per_hour_basis": "one measured whole solve (initialisation + all W rows, height caps on)",
        "solve_s": solve_s,
        "solve_kernel_s": ss["kernel_s"],
        "verified_first_residue": verified,
        "residues": residues,
        "moduli": mods,
        "fixture_checks": checked,
        "plain_k1": plain,
        "batched_vs_plain_residues_per_hour": (residues_per_hour / plain["residues_per_hour"]) if plain else None,
        "roofline": roof,
        "roofline_kernels": list(kern.values()) + ([capped] if capped else []),
        "counter_run": traffic,
        "hbm_copy": copy,
        "cpu_baseline": cpu,
        "plain_small": small,
        "shard_n27": shard,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
