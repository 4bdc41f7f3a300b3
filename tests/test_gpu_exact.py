"""The exact-run configuration on the GPU: the n=22 solve with every CRT modulus
the bound needs interleaved in one sweep, and solves whose moduli are split
over several sweeps.  Through the C ABI, bit-exact."""
from __future__ import annotations

import os
import re
import subprocess

import pytest

from conftest import PKG, load_golden

pytestmark = pytest.mark.gpu
CLI = PKG / "lib" / "ofp_solver"


def _fixtures(n: int) -> dict[int, int]:
    out = {}
    for name in ("reference_residues.json", "ggcount_residues.json", "ggcount_residues_large.json"):
        try:
            rows = load_golden(name)["residues"]
        except FileNotFoundError:
            continue
        for r in rows:
            if r["n"] == n:
                out[r["modulus"]] = r["residue"]
    return out


def test_exact_n22_fourteen_moduli():
    """n=22 with the 14 primes of the checkerboard-strip bound in one interleaved
    sweep: the first residue is the reference's (docs/results/
    groupbatch-interleaved-2026-09-05.json:1027), every residue an independent
    fixture holds matches, two more equal one-modulus solves, and the CRT integer
    lies below the bound and the modulus product."""
    import ofp
    from ofp.exact_safety import crt_pair
    from ofp.path_bound import primes_for_n, simple_path_upper_bound
    mods = primes_for_n(22)
    assert len(mods) == 14 and mods[0] == 4294967291
    with ofp.Context() as ctx:
        print("case n=22 k=14 whole interleaved solve", flush=True)
        got, st = ctx.solve(22, mods)
        assert st["rows"] == 23  # one sweep carried all 14 moduli
        assert got[0] == 1245069298
        fx = _fixtures(22)
        for m, r in fx.items():
            if m in mods:
                assert got[mods.index(m)] == r, m
        for j in (6, 13):
            print(f"case n=22 one-modulus solve of modulus {mods[j]}", flush=True)
            one, _ = ctx.solve(22, [mods[j]])
            assert one == [got[j]], j
    x, m = 0, 1
    for p, r in zip(mods, got):
        assert 0 <= r < p
        x, m = crt_pair(x, m, r, p)
    bound = simple_path_upper_bound(22)[0]
    assert m > bound >= x > 0
    # A007764(n) is even for n >= 1: reflection in the main diagonal maps the
    # corner-to-corner paths onto themselves and fixes none (it swaps the two
    # possible first edges), so they pair off.  A wrong residue would make the
    # CRT integer's parity a coin toss.
    assert x % 2 == 0


def test_solve_in_several_sweeps(gpu_ctx):
    """Moduli split over sweeps of at most 2, 3 and 5 give the residues of one
    sweep; the row count shows the split happened."""
    mods = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161, 4294967143]
    try:
        gpu_ctx.set_max_batch(0)
        want, st = gpu_ctx.solve(13, mods)
        assert st["rows"] == 14
        for g in (2, 3, 5):
            print(f"case max_batch={g}", flush=True)
            gpu_ctx.set_max_batch(g)
            got, st = gpu_ctx.solve(13, mods)
            assert got == want, g
            assert st["rows"] == 14 * -(-len(mods) // g)
    finally:
        gpu_ctx.set_max_batch(0)


def test_cli_runs_moduli_in_several_sweeps():
    """OFP_BATCH_MODULI caps the moduli per sweep: the CLI still prints one line per
    modulus in argument order with the same residues."""
    mods = ["4294967291", "4294966997", "2147483647", "65537", "1000003"]
    lines = {}
    for cap in ("0", "2"):
        env = dict(os.environ, OFP_BATCH_MODULI=cap)
        res = subprocess.run([str(CLI), "12", "16384", "14", "1", *mods], capture_output=True, text=True,
                             timeout=120, env=env)
        assert res.returncode == 0, res.stderr
        lines[cap] = [dict(re.findall(r"(\w+)=([^\s]+)", ln)) for ln in res.stdout.splitlines()
                      if ln.startswith("backend=")]
    assert [d["modulus"] for d in lines["2"]] == mods
    assert [d["residue"] for d in lines["2"]] == [d["residue"] for d in lines["0"]]
    assert lines["0"][0]["sweeps"] == "1" and lines["2"][0]["sweeps"] == "3"
    exact = int(load_golden("exact_counts.json")["counts"]["12"])
    assert [int(d["residue"]) for d in lines["2"]] == [exact % int(m) for m in mods]


def test_read_main_refuses_short_buffer(gpu_ctx):
    """A buffer shorter than the prepared moduli count is refused, not overrun."""
    import ofp
    mods = [4294967291, 4294966997, 2147483647]
    gpu_ctx.prepare(9, mods)
    gpu_ctx.row_sweep(10)
    with pytest.raises(ofp.OfpError) as e:
        gpu_ctx.read_main(0, 1)
    assert e.value.status == -1 and "3 moduli" in str(e.value)
    exact = int(load_golden("exact_counts.json")["counts"]["9"])
    assert gpu_ctx.read_main(0, 3) == [exact % m for m in mods]
