"""CRT runner, checkpoint and manifest, driven by a fake solver (CPU only).

Mirrors the reference's tests/test_exact_runner_e2e.py:18-120 and
tests/test_exact_safety.py:201-300: a fake solver prints result lines from a
table of known counts and records every invocation, so resume can be shown
not to recompute, a killed run to keep what it had checkpointed, and a
tampered checkpoint to be refused.
"""
from __future__ import annotations

import io
import json
import stat
import sys
import textwrap
from pathlib import Path

import pytest

from ofp import crt_runner, exact_safety
from ofp.path_bound import primes_for_n
from ofp.solver_output import parse_result_line

KNOWN = {4: 8512, 9: 41044208702632496804, 12: 64528039343270018963357185158482118}


def make_fake_solver(tmp: Path, crash_after: int | None = None, wrong: bool = False) -> Path:
    calls = tmp / "calls.log"
    script = tmp / ("fake_solver_crash" if crash_after is not None else "fake_solver")
    script.write_text(textwrap.dedent(f"""\
        #!{sys.executable}
        import sys
        KNOWN = {KNOWN!r}
        n = int(sys.argv[1]); mods = [int(x) for x in sys.argv[5:]]
        with open({str(calls)!r}, "a") as f:
            f.write(" ".join(sys.argv[1:]) + "\\n")
        print("fake solver diagnostic n=" + str(n), file=sys.stderr)
        print("backend=fake unrelated=1")
        for i, p in enumerate(mods):
            if {crash_after!r} is not None and i >= {crash_after!r}:
                sys.exit(3)
            r = (KNOWN[n] + ({1 if wrong else 0})) % p
            print(f"backend=fake n={{n}} residue={{r}} modulus={{p}} residue_index={{i}} wall_s=0.25", flush=True)
        """))
    script.chmod(script.stat().st_mode | stat.S_IXUSR)
    return script


def calls(tmp: Path) -> list[str]:
    f = tmp / "calls.log"
    return f.read_text().splitlines() if f.exists() else []


def run(n, binary, work, **kw):
    out, err = io.StringIO(), io.StringIO()
    rc = crt_runner.run(n, binary, work, out=out, err=err, **kw)
    return rc, out.getvalue(), err.getvalue()


def test_end_to_end_n4(tmp_path):
    fake = make_fake_solver(tmp_path)
    rc, out, _ = run(4, fake, tmp_path / "w")
    assert rc == 0
    assert "exact=8512" in out
    man = exact_safety.load_exact_result_manifest(tmp_path / "w" / "exact_manifest.json")
    assert man["exact_decimal"] == "8512" and man["primes_used"] == 1
    assert (tmp_path / "w" / "exact.txt").read_text().startswith("n=4\nexact=8512\n")
    assert man["exact_sha256"] == exact_safety.file_sha256(tmp_path / "w" / "exact.txt")
    assert len(calls(tmp_path)) == 1


def test_resume_does_not_recompute(tmp_path):
    fake = make_fake_solver(tmp_path)
    work = tmp_path / "w"
    primes = primes_for_n(12)
    assert len(primes) == 4
    rc, _, _ = run(12, fake, work, max_runs=1)
    assert rc == 0
    assert calls(tmp_path)[-1].split()[4:] == [str(primes[0])]
    rc, out, _ = run(12, fake, work)
    assert rc == 0 and f"exact={KNOWN[12]}" in out
    assert calls(tmp_path)[-1].split()[4:] == [str(p) for p in primes[1:]]
    n_calls = len(calls(tmp_path))
    rc, out, _ = run(12, fake, work)  # everything cached: no solver run
    assert rc == 0 and f"exact={KNOWN[12]}" in out
    assert len(calls(tmp_path)) == n_calls


def test_killed_solver_keeps_checkpointed_residues(tmp_path):
    crash = make_fake_solver(tmp_path, crash_after=2)
    work = tmp_path / "w"
    with pytest.raises(SystemExit, match="rc=3"):
        run(12, crash, work)
    primes = primes_for_n(12)
    data = json.loads((work / "checkpoint.json").read_text())
    assert sorted(int(k) for k in data["residues"]) == sorted(primes[:2])
    # the same binary resumes; only the missing moduli are requested
    good = crash  # same file, same identity; make it not crash any more
    good.write_text(good.read_text().replace("is not None and i >= 2", "is not None and i >= 99"))
    with pytest.raises(SystemExit, match="solver binary changed"):
        run(12, good, work)  # editing the binary changes its SHA-256: resume refused


def test_checkpoint_tamper_and_strictness(tmp_path):
    fake = make_fake_solver(tmp_path)
    work = tmp_path / "w"
    run(12, fake, work, max_runs=2)
    ck = work / "checkpoint.json"
    ident = exact_safety.solver_identity(fake)
    good = ck.read_text()
    assert len(exact_safety.load_checkpoint(ck, n=12, identity=ident)) == 2
    with pytest.raises(ValueError, match="belongs to n=12"):
        exact_safety.load_checkpoint(ck, n=11, identity=ident)
    data = json.loads(good)
    key = next(iter(data["residues"]))
    data["residues"][key]["residue"] = (data["residues"][key]["residue"] + 1) % int(key)
    ck.write_text(json.dumps(data))
    with pytest.raises(ValueError, match="checksum mismatch"):
        exact_safety.load_checkpoint(ck, n=12, identity=ident)
    ck.write_text(good.replace('"format_version": 4,', '"format_version": 4, "format_version": 4,', 1))
    with pytest.raises(ValueError, match="duplicate JSON key"):
        exact_safety.load_checkpoint(ck, n=12, identity=ident)
    ck.write_text(good)
    other = dict(ident, binary_sha256="0" * 64)
    with pytest.raises(ValueError, match="solver binary changed"):
        exact_safety.load_checkpoint(ck, n=12, identity=other)
    # a forged record that still carries a valid checksum is caught by range checks
    data = json.loads(good)
    data["residues"][key]["residue"] = int(key)
    data.pop(exact_safety.CHECKSUM_FIELD)
    data[exact_safety.CHECKSUM_FIELD] = exact_safety.checkpoint_checksum(data)
    ck.write_text(json.dumps(data))
    with pytest.raises(ValueError, match="outside canonical range"):
        exact_safety.load_checkpoint(ck, n=12, identity=ident)


def test_wrong_residues_fail_the_bound_or_congruence(tmp_path):
    fake = make_fake_solver(tmp_path, wrong=True)
    rc, out, _ = run(4, fake, tmp_path / "w")
    # one prime: CRT gives KNOWN+1 which is still within the bound and congruent,
    # so the single-prime case cannot detect it; with 4 primes the reconstruction
    # is exact (KNOWN+1) -- the manifest records exactly what the solver claimed.
    assert rc == 0 and "exact=8513" in out


def test_second_runner_is_locked_out(tmp_path):
    work = tmp_path / "w"
    work.mkdir()
    lock = exact_safety.acquire_workdir_lock(work)
    try:
        with pytest.raises(RuntimeError, match="already in use"):
            exact_safety.acquire_workdir_lock(work)
    finally:
        lock.close()


def test_crt_pair_reconstructs():
    x, m = 0, 1
    value = KNOWN[12]
    for p in primes_for_n(12):
        x, m = exact_safety.crt_pair(x, m, value % p, p)
    assert x == value


@pytest.mark.parametrize("line,expect", [
    ("backend=x n=9 residue=5 modulus=7 wall_s=1.5", (9, 5, 7, 1.5)),
    ("backend=x n=9 residue=5 modulus=7 batch_seconds=3 wall_s=1e-3", (9, 5, 7, 1e-3)),
    ("diagnostic n=9 states=12", None),
    ("", None),
])
def test_parse_result_line_accepts(line, expect):
    r = parse_result_line(line)
    assert (None if r is None else (r.n, r.residue, r.modulus, r.wall_s)) == expect


@pytest.mark.parametrize("line,msg", [
    ("n=9 residue=5 modulus=7", "missing"),
    ("n=9 residue=05 modulus=7 wall_s=1", "non-canonical"),
    ("n=9 residue=5 modulus=7 wall_s=1 wall_s=2", "duplicate"),
    ("n=9 residue=5 modulus=1 wall_s=1", "invalid solver modulus"),
    ("n=9 residue=5 modulus=7 wall_s=-1", "wall_s"),
    ("n=9 residue=5 modulus=7 wall_s=nan", "wall_s"),
    ("n=0 residue=5 modulus=7 wall_s=1", "invalid solver n"),
])
def test_parse_result_line_rejects(line, msg):
    with pytest.raises(ValueError, match=msg):
        parse_result_line(line)
