"""Serving loop: continuous batching of (n, modulus-set) solves on one GPU.

Requests arrive over time.  The scheduler keeps every admitted solve's
frontier state resident on the device as a job (ofp.Job: its own stream,
its own interleaved state) and

  * admits pending work whenever a job slot and HBM are free, using the
    state-size model (bytes = (main + blocked states) x padded k x 4, the same
    figure as ofp_plan) against a fixed budget (free HBM at start minus the
    reserve) -- the reference's admission test is its HBM budget check
    (row6crt20.cu:1751-1755);
  * batches at admission: all pending requests with the same n are merged into
    ONE job whose moduli are the union of theirs (duplicates shared), up to 16
    moduli (one cell launch), so their residues ride in one interleaved sweep;
  * launches each job with the tuned settings of its (n, k) (ofp.autotune's
    table; they are read when the rows are enqueued, so jobs of different
    sizes in flight each keep their own);
  * keeps jobs overlapping on the GPU: each job's W row sweeps are enqueued on
    its own stream as soon as it is admitted, the host only polls and retires;
  * retires a finished job: reads its residues, answers every request it
    carried, frees its memory, and admits more.

Each request's residues are bit-identical to a solo solve (same kernel, same
arithmetic, moduli are independent lanes of the interleaved vector); the GPU
test tests/test_gpu_serving.py checks exactly that.
"""
from __future__ import annotations

import random
import statistics
import time
from dataclasses import dataclass, field

from . import _lib, autotune
from .path_bound import PRIMES

MAX_MODULI_PER_JOB = 16


@dataclass
class Request:
    rid: int
    n: int
    moduli: list[int]
    arrival: float                      # seconds from the start of the stream
    done: float | None = None
    residues: list[int] | None = None
    job_moduli: int = 0                 # size of the merged job that served it

    @property
    def latency(self) -> float:
        return (self.done - self.arrival) if self.done is not None else float("nan")


@dataclass
class _Active:
    job: _lib.Job
    requests: list[Request]
    moduli: list[int]
    bytes: int
    started: float


def state_bytes(n: int, k: int, layout: int = 1) -> int:
    nm, nb = _lib.state_counts(n)
    pad = lambda x: (x + 63) // 64 * 64  # noqa: E731
    kp = (k + 1) // 2 * 2 if (layout == 1 and k > 1) else k
    return (pad(nm) + pad(nb)) * kp * 4


def poisson_stream(count: int, rate: float, seed: int, n_lo: int = 14, n_hi: int = 21,
                   k_lo: int = 1, k_hi: int = 8, pool: int = 16) -> list[Request]:
    """Seeded Poisson arrivals; n ~ U{n_lo..n_hi}; k ~ U{k_lo..k_hi} distinct
    moduli drawn from the first `pool` CRT primes."""
    rng = random.Random(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += rng.expovariate(rate)
        n = rng.randint(n_lo, n_hi)
        k = rng.randint(k_lo, k_hi)
        mods = sorted(rng.sample(PRIMES[:pool], k), reverse=True)
        out.append(Request(i, n, mods, t))
    return out


@dataclass
class Server:
    ctx: _lib.Context
    budget_bytes: int
    max_jobs: int = 8
    merge: bool = True
    tuned: bool = True
    pending: list[Request] = field(default_factory=list)
    active: list[_Active] = field(default_factory=list)
    used_bytes: int = 0
    jobs_started: int = 0
    peak_bytes: int = 0
    peak_jobs: int = 0

    def submit(self, req: Request) -> None:
        self.pending.append(req)

    def _admit(self, now: float) -> None:
        while self.pending and len(self.active) < self.max_jobs:
            head = self.pending[0]
            group = [head]
            mods = list(head.moduli)
            if self.merge:
                for r in self.pending[1:]:
                    if r.n != head.n:
                        continue
                    union = sorted(set(mods) | set(r.moduli), reverse=True)
                    if len(union) <= MAX_MODULI_PER_JOB:
                        group.append(r)
                        mods = union
            need = state_bytes(head.n, len(mods))
            if need > self.budget_bytes:
                raise RuntimeError(f"request n={head.n} k={len(mods)} needs {need} bytes > budget")
            if self.used_bytes + need > self.budget_bytes:
                return  # wait for memory: admission is FIFO on the head request
            if self.tuned:
                autotune.apply(self.ctx, head.n, len(mods))
            job = _lib.Job(self.ctx, head.n, mods)
            job.enqueue(head.n + 1)  # all W rows, asynchronously on the job's stream
            for r in group:
                self.pending.remove(r)
                r.job_moduli = len(mods)
            self.active.append(_Active(job, group, mods, need, now))
            self.used_bytes += need
            self.jobs_started += 1
            self.peak_bytes = max(self.peak_bytes, self.used_bytes)
            self.peak_jobs = max(self.peak_jobs, len(self.active))

    def _retire(self, now_fn) -> int:
        finished = 0
        for a in list(self.active):
            if not a.job.info().idle:
                continue
            res = a.job.result()
            t = now_fn()
            by_mod = dict(zip(a.moduli, res))
            for r in a.requests:
                r.residues = [by_mod[m] for m in r.moduli]
                r.done = t
            a.job.close()
            self.active.remove(a)
            self.used_bytes -= a.bytes
            finished += len(a.requests)
        return finished

    def run(self, requests: list[Request], poll_s: float = 0.0005, log=None) -> dict:
        """Replay a request stream in real time; returns the serving report."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0  # noqa: E731
        i = 0
        done = 0
        while done < len(reqs):
            t = now()
            while i < len(reqs) and reqs[i].arrival <= t:
                self.submit(reqs[i])
                i += 1
            self._admit(t)
            got = self._retire(now)
            done += got
            if got and log:
                log(f"serving: t={now():.2f}s done={done}/{len(reqs)} active={len(self.active)} "
                    f"pending={len(self.pending)} used={self.used_bytes / 2**30:.1f} GiB")
            if not got:
                nxt = reqs[i].arrival - now() if i < len(reqs) else poll_s
                time.sleep(max(0.0, min(poll_s, nxt)))
        wall = now()
        lat = sorted(r.latency for r in reqs)
        residues = sum(len(r.moduli) for r in reqs)

        def pct(q):
            return lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))]
        return {
            "requests": len(reqs),
            "residues": residues,
            "wall_s": wall,
            "residues_per_hour": 3600.0 * residues / wall,
            "latency_p50_s": pct(0.50),
            "latency_p99_s": pct(0.99),
            "latency_mean_s": statistics.fmean(lat),
            "jobs": self.jobs_started,
            "peak_active_jobs": self.peak_jobs,
            "peak_state_bytes": self.peak_bytes,
            "budget_bytes": self.budget_bytes,
        }


def default_budget(ctx: _lib.Context, reserve: int = 8 << 30) -> int:
    """Free HBM now minus a reserve (the reference's default reserve is
    clamp(total/32, 256 MiB, 8 GiB), fullmate_dropN_batch.cu:499)."""
    plan = ctx.plan(1, 1)
    return max(0, plan.free_bytes - reserve)


def main(argv=None) -> int:
    import argparse
    import json
    import sys
    ap = argparse.ArgumentParser(description="replay a Poisson stream of solves through the serving loop")
    ap.add_argument("--requests", type=int, default=40)
    ap.add_argument("--rate", type=float, default=2.0, help="mean arrivals per second")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--n-lo", type=int, default=14)
    ap.add_argument("--n-hi", type=int, default=21)
    ap.add_argument("--k-hi", type=int, default=8)
    ap.add_argument("--max-jobs", type=int, default=8)
    ap.add_argument("--no-merge", action="store_true")
    ap.add_argument("--untuned", action="store_true", help="run every job with the library's default settings")
    ap.add_argument("--check", action="store_true",
                    help="re-solve every request alone afterwards and require identical residues")
    a = ap.parse_args(argv)
    reqs = poisson_stream(a.requests, a.rate, a.seed, a.n_lo, a.n_hi, 1, a.k_hi)
    with _lib.Context() as ctx:
        srv = Server(ctx, default_budget(ctx), max_jobs=a.max_jobs, merge=not a.no_merge, tuned=not a.untuned)
        rep = srv.run(reqs, log=lambda m: print(m, file=sys.stderr, flush=True))
        rep["config"] = {"rate": a.rate, "seed": a.seed, "n": [a.n_lo, a.n_hi], "k": [1, a.k_hi],
                         "max_jobs": a.max_jobs, "merge": not a.no_merge, "tuned": not a.untuned}
        if a.check:
            bad = 0
            for r in reqs:
                solo, _ = ctx.solve(r.n, r.moduli)
                bad += solo != r.residues
            rep["solo_mismatches"] = bad
    print(json.dumps(rep), flush=True)
    return 0 if rep.get("solo_mismatches", 0) == 0 else 1


if __name__ == "__main__":
    raise SystemExit(main())
