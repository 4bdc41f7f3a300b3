#!/usr/bin/env python3
"""Benchmark: frontier transitions/s and residues/hour, batched vs one-residue,
on one MI355X, for the corner-to-corner simple-path count at n=22.

A *step* is one row sweep (cells W-1..1, W = n+1) of the interleaved frontier
state: every main and blocked word, all k CRT moduli, in HBM.  With the
defaults (--warmup 1 --steps 22) warmup + timed steps are the W = 23 rows of
one complete n=22 solve, so the residues are read afterwards and checked
against the reference's published value for the first prime.

Workload: n=22, k = number of primes the checkerboard-strip bound needs (14),
the largest primes below 2^32 -- an exact-count (CRT) run.  Inputs: the
solve's own initial state (synthetic in the sense that nothing is read from
disk).  The first rows and the last two rows of a solve sweep only the words
of small height (height caps; the answer is unchanged), so rows differ in
cost: residues/hour comes from the time of the whole solve (initialisation,
warmup row and timed rows; the state memory is allocated before), and `value`
counts every row at its full
algorithmic work W-1 cells x S states x k moduli.  Also measured in the same
process:
  * plain: the same solve with k=1 (one residue at a time), timed whole
    -> residues/hour of the plain form;
  * roofline of the cell kernel: declared algorithmic bytes per launch
    (8 B per state per modulus: read + write of one uint32 residue) divided by
    the average duration of the full-state launches (rows without capped
    cells), from HIP events on the library's stream;
  * plain_small: whole one-modulus solves at n=12 and n=18 (the small plain-form
    configurations; median of three, the n=18 residue checked against the
    reference's published value);
  * cpu_baseline: the CPU oracle's pull-form step timed on the host cores on
    a bounded sample of the same n=22 cell step, and its whole n=12 solve
    (median of three, residue compared with the GPU's).

Prints ONE JSON line on stdout.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "oneesan-gpu_amd"))
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
GOLDEN_FIRST = {22: (4294967291, 1245069298), 21: (4294967291, 998035516),
                20: (4294967291, 2308006916), 18: (4294967291, 503411004)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _cpu_model() -> str | None:
    """The host CPU's model name (first 'model name' line of /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(n: int, k: int, seconds: float) -> dict:
    """Time the CPU oracle's pull-form cell step on a bounded sample of the
    n-sized state (consecutive target windows of a middle cell), at one
    modulus; rate = target states processed per second."""
    import ctypes

    import numpy as np

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


def cpu_solve_n12(mod: int) -> dict:
    """The CPU oracle's whole n=12 solve (pull form, one modulus): one warm-up,
    then the median of three."""
    from oracle import binding
    binding.solve(12, mod)
    times, res = [], None
    for _ in range(3):
        t0 = time.perf_counter()
        res = binding.solve(12, mod)
        times.append(time.perf_counter() - t0)
    med = sorted(times)[1]
    T = 13 * 12 * 98123
    return {"solve_s": med, "residue": res, "transitions_per_s": T / med,
            "sample": "oracle whole n=12 solve (oracle/gridfp_oracle.c or_solve, pull form), median of 3"}


def copy_calibration(nbytes: int) -> int:
    """One device-to-device copy of nbytes (HIP runtime blit kernel), so the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters can be scaled: a copy of B bytes
    must read and write B bytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(nbytes)) != 0:
            log("copy calibration: hipMalloc failed")
            return 1
    hip.hipMemset(src, 1, ctypes.c_size_t(nbytes))
    hip.hipDeviceSynchronize()
    t0 = time.perf_counter()
    rc = hip.hipMemcpy(dst, src, ctypes.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    el = time.perf_counter() - t0
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copy_bytes": nbytes, "seconds": el, "GBps_read_plus_write": 2 * nbytes / el / 1e9,
                      "rc": rc}), flush=True)
    return 0 if rc == 0 else 1


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=22, help="timed row sweeps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed row sweeps before timing")
    ap.add_argument("--n", type=int, default=22)
    ap.add_argument("--k", type=int, default=0, help="moduli (0 = as many as the CRT bound needs)")
    ap.add_argument("--no-plain", action="store_true", help="skip the k=1 plain-form solve")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-shard", action="store_true", help="skip the n=27 one-GPU-share measurement")
    ap.add_argument("--no-small", action="store_true",
                    help="skip the whole plain-form solves at n=12 and n=18 (k=1)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark measures one GPU: the frontier state of one solve "
                                   "lives in one GPU's HBM; run with --gpus 1"}), flush=True)
        return 2

    import ofp
    from ofp.path_bound import PRIMES, primes_for_n

    n = args.n
    W = n + 1
    mods = primes_for_n(n) if args.k == 0 else PRIMES[:args.k]
    k = len(mods)
    nm, nb = ofp.state_counts(n)
    S = nm + nb
    log(f"bench: n={n} W={W} k={k} states={S} ({nm} main + {nb} blocked)")

    from ofp import autotune
    ctx = ofp.Context(device=0)
    tuned = autotune.apply(ctx, n, k)
    log(f"bench: launch config {tuned}")
    plan = ctx.plan(n, k)
    log(f"bench: state {plan.state_bytes / 2**30:.1f} GiB, free {plan.free_bytes / 2**30:.1f} GiB")
    # first prepare allocates the state (the operating system may clear freshly mapped
    # device pages first: allocation, not solve time); the timed solve starts from the
    # resident state
    ctx.prepare(n, mods)
    tw0 = time.perf_counter()
    ctx.prepare(n, mods)
    if args.warmup:
        ctx.row_sweep(args.warmup)
        log(f"bench: warmup {args.warmup} rows done")
    warm_s = time.perf_counter() - tw0
    # The library synchronises its stream before returning from row_sweep, so the
    # timed region is bracketed by device synchronisation on both sides.
    t0 = time.perf_counter()
    st = ctx.row_sweep(args.steps)
    t1 = time.perf_counter()
    el = t1 - t0
    log(f"bench: {args.steps} rows in {el:.3f} s (kernel {st['kernel_s']:.3f} s, {st['launches']} launches)")
    verified = None
    residues = None
    if args.warmup + args.steps == W:
        residues = ctx.read_main(0, k)
        if n in GOLDEN_FIRST and GOLDEN_FIRST[n][0] == mods[0]:
            verified = residues[0] == GOLDEN_FIRST[n][1]
        log(f"bench: residues {residues} verified_first={verified}")

    transitions = st["transitions"]
    value = transitions / el
    ms_per_step = el / args.steps * 1e3
    whole = args.warmup + args.steps == W
    # a whole solve: initialisation + warmup + timed rows; otherwise extrapolated
    solve_s = (warm_s + el) if whole else el / args.steps * W
    residues_per_hour = 3600.0 * k / solve_s
    if st["full_launches"]:
        avg_launch_s = st["full_kernel_s"] / st["full_launches"]
        launch_basis = f"{st['full_launches']} full-state launches (rows without capped cells)"
    else:
        avg_launch_s = st["kernel_s"] / st["launches"]
        launch_basis = f"all {st['launches']} launches (no row of the timed region ran uncapped)"
    bytes_per_launch = 8.0 * S * min(k, 16)  # one launch carries up to 16 moduli
    achieved = bytes_per_launch / avg_launch_s / 1e9
    # HBM bytes per full-state launch as measured with rocprofv3 counters for this
    # exact launch configuration (recorded with the tuned entry), else unknown
    traffic, traffic_source = None, None
    if tuned and tuned.get("hbm_bytes_per_full_launch"):
        traffic = float(tuned["hbm_bytes_per_full_launch"])
        traffic_source = tuned.get("hbm_bytes_source")

    # plain one-residue-at-a-time form (k=1): one whole solve (state memory already held)
    plain = None
    if not args.no_plain:
        autotune.apply(ctx, n, 1)
        ctx.prepare(n, [mods[0]])
        tp0 = time.perf_counter()
        ctx.prepare(n, [mods[0]])
        sp = ctx.row_sweep(W)
        tp = time.perf_counter() - tp0
        pres = ctx.read_main(0, 1)
        plain = {"solve_s": tp, "residues_per_hour": 3600.0 / tp,
                 "transitions_per_s": sp["transitions"] / tp,
                 "full_row_ms": (sp["full_kernel_s"] / sp["full_launches"] * (W - 1) * 1e3
                                 if sp["full_launches"] else None),
                 "residue": pres[0], "residue_matches_batched": residues is not None and pres[0] == residues[0]}
        if sp["full_launches"]:
            # the same declared bytes as the batched roofline, at one modulus; this form
            # is bound by its integer index work, not by HBM
            pl_s = sp["full_kernel_s"] / sp["full_launches"]
            plain.update({"avg_launch_ms": pl_s * 1e3, "declared_GBps": 8.0 * S / pl_s / 1e9,
                          "hbm_frac": 8.0 * S / pl_s / 1e9 / HBM_PEAK_GBS})
        log(f"bench: plain k=1 solve {tp:.2f} s -> {plain['residues_per_hour']:.0f} residues/h "
            f"(residue {pres[0]})")
    ctx.close()

    # whole plain-form solves of the small configurations (n=12, n=18, one modulus):
    # one untimed solve (graph capture), then the median of three
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
        with ofp.Context(device=0) as sctx:
            autotune.apply(sctx, 22, 4)
            shard = shard_mod.measure(sctx)
        log(f"bench: n=27 share window {shard['window_bytes'] / 2**30:.1f} GiB, "
            f"{shard['avg_launch_ms']:.1f} ms/cell, {shard['roofline']['frac']:.3f} of HBM peak")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(n, k, args.cpu_seconds)
        log(f"bench: cpu baseline {cpu['value']:.3e} transitions/s on {cpu['cores']} threads")
        cpu["n12"] = cpu_solve_n12(PRIMES[0])
        if small:
            cpu["n12"]["residue_matches_gpu"] = cpu["n12"]["residue"] == small["n12"]["residue"]

    out = {
        "metric": "frontier transitions/s & residues/hour, batched vs 1-residue, n=22, bit-exact",
        "value": value,
        "unit": "transitions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (the solve's own initial frontier; nothing loaded)",
        "config": {"workload": f"exact CRT count, n={n} ({W}x{W} vertices), {k} moduli interleaved, "
                               f"one row sweep per step", "n": n, "moduli": k, "states": S,
                   "launch": tuned,
                   "main_states": nm, "blocked_states": nb, "state_bytes": plan.state_bytes,
                   "parallelism": "single GPU"},
        "residues_per_hour": residues_per_hour,
        "solve_s": solve_s,
        "verified_first_residue": verified,
        "residues": residues,
        "plain_k1": plain,
        "batched_vs_plain_residues_per_hour": (residues_per_hour / plain["residues_per_hour"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_source,
                     "kernel": "cell_kernel", "bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3, "launch_basis": launch_basis,
                     "capped_kernel_s": st["capped_kernel_s"], "capped_launches": st["capped_launches"]},
        "cpu_baseline": cpu,
        "plain_small": small,
        "shard_n27": shard,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
