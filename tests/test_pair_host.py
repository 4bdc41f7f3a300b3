"""The pair pass's per-class body (oneesan-gpu_amd/csrc/ofp_pair_impl.h), built
for the host and run over every class of every (p, d), against the CPU oracle's
two-cell windows: n = 2..9, all pairs of cells, blocked input on and off, 32-bit
ranks (and 64-bit ranks at n = 3, 6, 9), three moduli including 2.  No GPU."""
from __future__ import annotations

import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = ROOT / "tests" / "native" / "pair_host_check.cpp"
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory, oracle):
    if not shutil.which(HIPCC) and not (ROOT / HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("pair") / "pair_host_check"
    lib = ROOT / "oracle"
    cmd = [HIPCC, "-O1", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(SRC), "-o", str(out), f"-L{lib}", "-loracle", f"-Wl,-rpath,{lib}"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out


def test_pair_body_equals_two_oracle_cells(checker):
    res = subprocess.run([str(checker), "9"], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert res.stdout.strip().endswith("ALL OK")
    assert res.stdout.count("\nok ") >= 80
