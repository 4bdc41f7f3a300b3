"""The drop-in CLI and the CRT runner on the GPU (the reference process contract)."""
from __future__ import annotations

import io
import re
import subprocess

import pytest

from conftest import PKG, load_golden

pytestmark = pytest.mark.gpu
CLI = PKG / "lib" / "ofp_solver"


def test_cli_result_lines_n9():
    """One line per modulus, in argument order, n/residue/modulus/wall_s once each and
    in order (tests/test_solver_contract.py:38-49); values as the reference's n=9
    runtime test expects (scripts/test/frontier-runtime.py:19-41)."""
    res = subprocess.run([str(CLI), "9", "16384", "14", "1", "4294967291", "4294966997"],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    lines = [ln for ln in res.stdout.splitlines() if ln.startswith("backend=")]
    assert len(lines) == 2
    want = {4294967291: 2674633373, 4294966997: 3332982389}
    for ln, p in zip(lines, want):
        toks = [t.split("=")[0] for t in ln.split()]
        for key in ("n", "residue", "modulus", "wall_s"):
            assert toks.count(key) == 1
        assert [toks.index(k) for k in ("n", "residue", "modulus", "wall_s")] == sorted(
            toks.index(k) for k in ("n", "residue", "modulus", "wall_s"))
        kv = dict(re.findall(r"(\w+)=([^\s]+)", ln))
        assert int(kv["n"]) == 9 and int(kv["modulus"]) == p and int(kv["residue"]) == want[p]
        assert ln.count("wall_s=") == 1 and ln.count(" n=") == 1


def test_cli_default_modulus_n12():
    res = subprocess.run([str(CLI), "12"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    exact = int(load_golden("exact_counts.json")["counts"]["12"])
    kv = dict(re.findall(r"(\w+)=([^\s]+)", res.stdout.strip().splitlines()[-1]))
    assert int(kv["modulus"]) == 4294967291 and int(kv["residue"]) == exact % 4294967291


def test_crt_runner_exact_n12(tmp_path):
    """Exact CRT count through the runner and the real binary; resume is a no-op."""
    from ofp import crt_runner, exact_safety
    out, err = io.StringIO(), io.StringIO()
    rc = crt_runner.run(12, CLI, tmp_path / "w", out=out, err=err)
    assert rc == 0, err.getvalue()
    exact = int(load_golden("exact_counts.json")["counts"]["12"])
    assert f"exact={exact}" in out.getvalue()
    man = exact_safety.load_exact_result_manifest(tmp_path / "w" / "exact_manifest.json")
    assert int(man["exact_decimal"]) == exact and man["primes_used"] == 4
    out2 = io.StringIO()
    assert crt_runner.run(12, CLI, tmp_path / "w", out=out2, err=io.StringIO()) == 0
    assert "batch run" not in out2.getvalue()
    # the result verifies from first principles, with the build's provenance sidecar
    from ofp import verify
    assert man["build_provenance_file"] == "build_provenance.json"
    assert verify.verify(tmp_path / "w" / "exact_manifest.json", binary=CLI)["n"] == 12
