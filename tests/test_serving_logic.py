"""Scheduler logic of the serving loop without a GPU: a fake job backend that
finishes after a few polls, so admission, merging and memory accounting are
checked on the CPU."""
from __future__ import annotations

import ofp._lib as lib
from ofp import serving
from ofp.path_bound import PRIMES


class FakeJob:
    live = 0
    peak = 0

    def __init__(self, ctx, n, moduli):
        self.n, self.moduli, self.polls, self.rows = n, list(moduli), 0, 0
        FakeJob.live += 1
        FakeJob.peak = max(FakeJob.peak, FakeJob.live)

    def enqueue(self, rows):
        self.rows += rows

    def info(self):
        self.polls += 1

        class I:
            pass
        i = I()
        i.idle = int(self.polls > 2 and self.rows == self.n + 1)
        return i

    def result(self):
        return [(self.n * 1000003 + p) % p for p in self.moduli]

    def close(self):
        FakeJob.live -= 1


def test_admission_merge_and_budget(monkeypatch):
    monkeypatch.setattr(lib, "Job", FakeJob)
    monkeypatch.setattr(serving._lib, "Job", FakeJob)
    reqs = [serving.Request(i, n, mods, 0.0) for i, (n, mods) in enumerate([
        (10, PRIMES[:3]), (11, PRIMES[2:4]), (10, PRIMES[1:5]), (10, PRIMES[:1]), (12, PRIMES[:8]),
        (10, PRIMES[10:16]), (10, PRIMES[:16])])]
    budget = serving.state_bytes(12, 8) + serving.state_bytes(10, 16) + 1
    srv = serving.Server(ctx=None, budget_bytes=budget, max_jobs=2, merge=True, tuned=False)
    rep = srv.run(reqs, poll_s=0.0)
    assert rep["requests"] == len(reqs)
    assert srv.peak_bytes <= budget and FakeJob.peak <= 2
    for r in reqs:
        assert r.residues == [(r.n * 1000003 + p) % p for p in r.moduli]
    # all five n=10 requests are pending together and their moduli union is
    # PRIMES[:16], so they ride in one 16-modulus job; n=11 and n=12 get their own
    assert all(r.job_moduli == 16 for r in reqs if r.n == 10)
    assert rep["jobs"] == 3


def test_poisson_stream_is_seeded():
    a = serving.poisson_stream(20, 2.0, 5)
    b = serving.poisson_stream(20, 2.0, 5)
    assert [(r.n, r.moduli, r.arrival) for r in a] == [(r.n, r.moduli, r.arrival) for r in b]
    assert all(14 <= r.n <= 21 and 1 <= len(r.moduli) <= 8 for r in a)
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
