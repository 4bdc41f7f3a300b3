"""Exact-result verifier and build provenance sidecar (CPU only).

Mirrors the reference's tests of scripts/tools/verify_exact_result.py and
scripts/solve/build_provenance.py: a result produced by the CRT runner (driven
by a fake solver) verifies; each kind of tampering is refused with its own
message; the sidecar written at build time matches the built solver.
"""
from __future__ import annotations

import json
import shutil
from pathlib import Path

import pytest

from ofp import exact_safety, path_bound, provenance, verify
from test_exact_runner import make_fake_solver, run

ROOT = Path(__file__).resolve().parents[1]


def _result(tmp_path, n=12, with_provenance=False):
    fake = make_fake_solver(tmp_path)
    if with_provenance:
        lib = tmp_path / "lib"
        lib.mkdir()
        src = tmp_path / "cli.c"
        src.write_text('#include "hdr.h"\nint main(){}\n')
        (tmp_path / "hdr.h").write_text("/* header */\n")
        provenance.create(fake, source=src, compiler=Path(shutil.which("gcc")), compile_args=["-O2"],
                          auxiliary=[("library", src)], root=tmp_path)
    work = tmp_path / "w"
    rc, _, _ = run(n, fake, work)
    assert rc == 0
    return work / "exact_manifest.json", fake


def _rewrite(manifest: Path, **changes):
    data = json.loads(manifest.read_text())
    data.update(changes)
    data[exact_safety.RESULT_CHECKSUM_FIELD] = exact_safety.result_manifest_checksum(data)
    manifest.write_text(json.dumps(data, indent=2, sort_keys=True) + "\n")


def test_independent_bound_matches_runner_bound():
    for n in (1, 2, 5, 9, 12, 16, 22, 27):
        assert verify.path_bound(n) == path_bound.simple_path_upper_bound(n)
    assert verify.strip_count(3, 4) == path_bound.strip_count(3, 4)


def test_valid_result_verifies(tmp_path, capsys):
    man, fake = _result(tmp_path)
    assert verify.main([str(man), "--binary", str(fake)]) == 0
    out = capsys.readouterr().out
    assert "valid=1" in out and "exact=64528039343270018963357185158482118" in out


@pytest.mark.parametrize("edit,msg", [
    (lambda m: _rewrite(m, exact_decimal=str(int(json.loads(m.read_text())["exact_decimal"]) + 1)),
     "exact value mismatch"),
    (lambda m: _rewrite(m, path_bound_decimal="12345"), "path bound mismatch"),
    (lambda m: _rewrite(m, solver_wall_s_sum=99.0), "wall-time sum mismatch"),
    (lambda m: _rewrite(m, primes_used=7), "primes_used mismatch"),
    (lambda m: _rewrite(m, congruences=json.loads(m.read_text())["congruences"][:-1],
                        primes_used=len(json.loads(m.read_text())["congruences"]) - 1), "does not exceed"),
    (lambda m: m.write_text(m.read_text().replace('"n": 12', '"n": 13')), "checksum mismatch"),
    (lambda m: (m.parent / "exact.txt").write_text("n=12\nexact=1\n"), "exact file SHA-256 mismatch"),
    (lambda m: (m.parent / "checkpoint.json").write_text("{}"), "checkpoint SHA-256 mismatch"),
])
def test_tampering_is_refused(tmp_path, edit, msg):
    man, _ = _result(tmp_path)
    edit(man)
    with pytest.raises(ValueError, match=msg):
        verify.verify(man)


def test_wrong_binary_is_refused(tmp_path):
    man, _ = _result(tmp_path)
    other = tmp_path / "other"
    other.write_text("not the solver")
    with pytest.raises(ValueError, match="binary SHA-256 mismatch"):
        verify.verify(man, binary=other)


def test_provenance_recorded_and_checked(tmp_path):
    man, fake = _result(tmp_path, with_provenance=True)
    data = json.loads(man.read_text())
    assert data["build_provenance_file"] == "build_provenance.json"
    pf = man.parent / "build_provenance.json"
    assert data["build_provenance_sha256"] == exact_safety.file_sha256(pf)
    assert verify.verify(man, binary=fake)["n"] == 12
    # a sidecar whose own checksum no longer matches is refused, even with a fixed-up manifest hash
    p = json.loads(pf.read_text())
    p["compile_args"] = ["-O0"]
    pf.write_text(json.dumps(p, indent=2, sort_keys=True) + "\n")
    _rewrite(man, build_provenance_sha256=exact_safety.file_sha256(pf))
    with pytest.raises(ValueError, match="build provenance SHA-256 checksum mismatch"):
        verify.verify(man)


def test_runner_refuses_sidecar_of_another_binary(tmp_path):
    fake = make_fake_solver(tmp_path)
    src = tmp_path / "cli.c"
    src.write_text("int main(){}\n")
    provenance.create(fake, source=src, compiler=Path(shutil.which("gcc")), compile_args=[],
                      auxiliary=[], root=tmp_path)
    fake.write_text(fake.read_text() + "\n# changed after the sidecar was written\n")
    with pytest.raises(SystemExit, match="build provenance does not match"):
        run(4, fake, tmp_path / "w")


def test_sidecar_of_the_built_solver():
    """build.py writes lib/ofp_solver.provenance.json; it must match the built CLI and this checkout."""
    lib = ROOT / "oneesan-gpu_amd" / "lib"
    side = lib / ("ofp_solver" + provenance.SIDECAR_SUFFIX)
    if not side.exists():
        pytest.skip("solver not built")
    d = provenance.load(side, binary=lib / "ofp_solver")
    assert d["format"] == provenance.FORMAT_V2
    roles = {r["role"] for r in d["auxiliary_dependencies"]}
    assert {"library", "library-source-cell", "library-source-api", "abi-header"} <= roles
    assert d["source"] == "oneesan-gpu_amd/csrc/ofp_cli.cpp"
