"""The benchmark's output contract on a small solve (bench.py as a child process).

One JSON line on stdout with the keys the benchmark contract names, a roofline
object with achieved / peak / frac, the residue of a whole n=12 solve equal to
the CPU oracle's, and `--gpus 2` refused with a message.
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _run(args, timeout=240):
    env = dict(os.environ)
    env["PYTHONPATH"] = os.pathsep.join([str(ROOT / "oneesan-gpu_amd"), str(ROOT)])
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True,
                          timeout=timeout, env=env, cwd=str(ROOT))


@pytest.mark.gpu
def test_bench_json_line_small_solve(oracle):
    print("case bench n=12 k=3 whole solve", flush=True)
    r = _run(["--n", "12", "--k", "3", "--warmup", "1", "--steps", "12", "--no-shard", "--no-small",
              "--cpu-seconds", "1"])
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout
    out = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better",
                "scaling", "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in out, key
    assert out["n_gpus"] == 1 and out["steps"] == 12 and out["warmup"] == 1
    assert out["value"] > 0 and out["dtype"] == "u32"
    roof = out["roofline"]
    assert roof["bound"] == "hbm" and roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-9
    assert len(out["residues"]) == 3
    from ofp.path_bound import PRIMES
    assert out["residues"] == [oracle.solve(12, p) for p in PRIMES[:3]]
    assert out["plain_k1"]["residue"] == out["residues"][0]
    assert out["cpu_baseline"]["cores"] >= 1 and out["cpu_baseline"]["n12"]["residue"] == out["residues"][0]


def test_bench_refuses_several_gpus():
    """Runs without a GPU: the refusal comes before any device work."""
    print("case bench --gpus 2", flush=True)
    r = _run(["--gpus", "2"], timeout=60)
    assert r.returncode != 0
    assert "one GPU" in r.stdout
