"""Write tests/golden/ggcount_residues_large.json: A007764(n) mod p for n=20..22
at CRT primes beyond the ones the reference publishes, from the reference's
third_party ggcount (Iwashita et al. 2013, MIT) built with its own OpenMP loops
by oracle/ref.mk into oracle/_ref/ggcount_omp.

Each case is appended to the fixture as soon as it finishes, so an interrupted
run keeps what it computed; cases already in the file are not recomputed.
Run in the container that has /root/reference, from the repository root:
    make -C oracle -f ref.mk && OMP_NUM_THREADS=6 python tests/golden/make_ggcount_large.py
"""
import json
import os
import re
import subprocess
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
GG = ROOT / "oracle" / "_ref" / "ggcount_omp"
OUT = ROOT / "tests" / "golden" / "ggcount_residues_large.json"
# n=20/21 at the second CRT prime (the reference publishes the first and
# 4294966997); n=22 at the second prime, then at the first prime, whose only
# published residue (1245069298) came from the reference's own GPU run.
CASES = [(20, 4294967279), (21, 4294967279), (22, 4294967279), (22, 4294967291), (22, 4294967231)]


def load() -> dict:
    if OUT.exists():
        return json.loads(OUT.read_text())
    return {"_about": "A007764(n) mod p from the reference's third_party/ggcount/src/ggcount.cpp "
                      "(independent frontier DP, OpenMP build by oracle/ref.mk); "
                      "command: ggcount_omp n %p",
            "residues": []}


def main() -> None:
    data = load()
    done = {(r["n"], r["modulus"]) for r in data["residues"]}
    for n, p in CASES:
        if (n, p) in done:
            continue
        t = time.time()
        res = subprocess.run([str(GG), str(n), f"%{p}"], capture_output=True, text=True, check=True)
        m = re.search(r"^(\d+) \(mod (\d+)\)$", res.stdout.strip().splitlines()[-1])
        assert m and int(m.group(2)) == p, res.stdout
        data["residues"].append({"n": n, "modulus": p, "residue": int(m.group(1)),
                                 "threads": int(os.environ.get("OMP_NUM_THREADS", "0")),
                                 "seconds": round(time.time() - t, 1)})
        OUT.write_text(json.dumps(data, indent=1) + "\n")
        print(n, p, m.group(1), round(time.time() - t, 1), flush=True)


if __name__ == "__main__":
    main()
