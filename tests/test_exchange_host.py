"""The exchange's index math (oneesan-gpu_amd/csrc/ofp_exchange_impl.h), built
for the host: every word of every split at n <= 10 goes from its LOW group
and LOW rank back to itself and to its HIGH group and HIGH rank (both counted
by enumeration), and at n = 22 and 27 random words agree with their groups'
own completion tables.  No GPU."""
from __future__ import annotations

import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = ROOT / "tests" / "native" / "exchange_host_check.cpp"
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory, oracle):
    if not shutil.which(HIPCC) and not (ROOT / HIPCC).exists():
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("xchg") / "exchange_host_check"
    lib = ROOT / "oracle"
    cmd = [HIPCC, "-O1", "-std=c++17", "-x", "c++", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
           str(SRC), "-o", str(out), f"-L{lib}", "-loracle", f"-Wl,-rpath,{lib}"]
    subprocess.run(cmd, check=True, capture_output=True, timeout=600)
    return out


def _run(checker, *args):
    res = subprocess.run([str(checker), *map(str, args)], capture_output=True, text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-2000:] + res.stderr[-2000:]
    assert res.stdout.strip().endswith("ALL OK")
    return res.stdout


def test_every_word_of_every_split(checker):
    out = _run(checker, "small", 10)
    assert "cases 90" in out


@pytest.mark.parametrize("n,low", [(22, 11), (26, 13), (27, 14), (27, 15)])
def test_random_words_at_full_size(checker, n, low):
    out = _run(checker, "big", n, low, 5000, n * 100 + low)
    assert out.count("ok big") == 2
