"""Serving loop on the GPU: overlapping jobs, merged moduli, memory admission.
Every request's residues must be bit-identical to a solo solve of the same
(n, moduli) and to the golden counts."""
from __future__ import annotations

import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu


def _golden(n, p):
    ex = load_golden("exact_counts.json")["counts"]
    if str(n) in ex:
        return int(ex[str(n)]) % p
    try:
        rows = load_golden("ggcount_residues.json")["residues"]
    except FileNotFoundError:
        return None
    for r in rows:
        if r["n"] == n and r["modulus"] == p:
            return r["residue"]
    return None


@pytest.mark.parametrize("merge,max_jobs,budget_gib", [(True, 4, 64), (False, 6, 64), (True, 8, 0.02)])
def test_served_equals_solo(gpu_ctx, merge, max_jobs, budget_gib):
    """Served with the tuned settings of each job's (n, k), in a context of its own;
    the solo solves run with the library's defaults."""
    import ofp
    from ofp import serving
    reqs = serving.poisson_stream(16, rate=200.0, seed=7 + max_jobs, n_lo=8, n_hi=14, k_lo=1, k_hi=5)
    budget = int(budget_gib * 2**30)
    print(f"case merge={merge} max_jobs={max_jobs} budget={budget}", flush=True)
    with ofp.Context() as sctx:
        srv = serving.Server(sctx, budget, max_jobs=max_jobs, merge=merge)
        rep = srv.run(reqs)
    assert rep["requests"] == 16
    assert srv.peak_bytes <= budget
    if merge:
        assert rep["jobs"] <= 16
    for r in reqs:
        solo, _ = gpu_ctx.solve(r.n, r.moduli)
        assert r.residues == solo, (r.rid, r.n, r.moduli)
        for p, res in zip(r.moduli, r.residues):
            g = _golden(r.n, p)
            if g is not None:
                assert res == g, (r.n, p)
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0


def test_jobs_overlap_and_results(gpu_ctx):
    """Several jobs enqueued back to back on their own streams all finish correctly."""
    import ofp
    primes = [4294967291, 4294966997, 2147483647]
    jobs = [ofp.Job(gpu_ctx, n, primes[: 1 + n % 3]) for n in (10, 11, 12, 13, 9, 14)]
    for j in jobs:
        j.enqueue(j.n + 1)
    for j in jobs:
        solo, _ = gpu_ctx.solve(j.n, j.moduli)
        assert j.result() == solo
        inf = j.info()
        assert inf.idle == 1 and inf.rows_enqueued == inf.rows_total == j.n + 1
        j.close()


def _fixture_residue(n, p):
    """Independent residues the fixtures hold for n >= 20 (reference, ggcount)."""
    for name in ("reference_residues.json", "ggcount_residues_large.json"):
        try:
            rows = load_golden(name)["residues"]
        except FileNotFoundError:
            continue
        for r in rows:
            if r["n"] == n and r["modulus"] == p:
                return r["residue"]
    return None


def test_served_stream_at_config_size():
    """The serving configuration at its stated size: a seeded Poisson stream (mean
    2 requests/s) of 10 solves with n in 14..21 and 1..8 moduli from the CRT prime
    list (this seed includes n=21 with 8 moduli, a 31.6 GB state), admitted against
    the free HBM.  Every request's residues equal its solo solve and every residue a
    fixture holds.  A context of its own, so that the budget is the HBM actually free."""
    import ofp
    from ofp import serving
    reqs = serving.poisson_stream(10, rate=2.0, seed=2, n_lo=14, n_hi=21, k_lo=1, k_hi=8)
    assert any(r.n == 21 and len(r.moduli) == 8 for r in reqs)
    with ofp.Context() as ctx:
        budget = serving.default_budget(ctx)
        srv = serving.Server(ctx, budget, max_jobs=8, merge=True)
        print(f"case config stream: {[(r.n, len(r.moduli)) for r in reqs]} budget={budget}", flush=True)
        rep = srv.run(reqs)
        assert rep["requests"] == 10 and srv.peak_bytes <= budget
        assert rep["peak_active_jobs"] >= 2  # requests overlapped on the device
        for r in reqs:
            print(f"case solo n={r.n} k={len(r.moduli)}", flush=True)
            solo, _ = ctx.solve(r.n, r.moduli)
            assert r.residues == solo, (r.rid, r.n, r.moduli)
            for p, res in zip(r.moduli, r.residues):
                g = _golden(r.n, p)
                if g is None:
                    g = _fixture_residue(r.n, p)
                if g is not None:
                    assert res == g, (r.n, p)
    assert rep["latency_p99_s"] >= rep["latency_p50_s"] > 0
