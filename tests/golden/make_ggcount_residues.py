"""Write tests/golden/ggcount_residues.json: A007764(n) mod p for n=10..19 from
the reference's third_party ggcount (Iwashita et al. 2013, MIT), built from its
own single source file by oracle/ref.mk into oracle/_ref/ggcount_single.
Run in the container that has /root/reference, from the repository root:
    make -C oracle -f ref.mk && python tests/golden/make_ggcount_residues.py
"""
import json
import re
import subprocess
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
GG = ROOT / "oracle" / "_ref" / "ggcount_single"
CASES = [(n, p) for n in range(10, 19) for p in (4294967291, 4294966997, 2147483647)]
CASES += [(19, 4294967291)]
out = {"_about": "A007764(n) mod p from the reference's third_party/ggcount/src/ggcount.cpp "
                 "(independent frontier DP, built by oracle/ref.mk); command: ggcount_single n %p",
       "residues": []}
for n, p in CASES:
    t = time.time()
    res = subprocess.run([str(GG), str(n), f"%{p}"], capture_output=True, text=True, check=True)
    m = re.search(r"^(\d+) \(mod (\d+)\)$", res.stdout.strip().splitlines()[-1])
    assert m and int(m.group(2)) == p, res.stdout
    out["residues"].append({"n": n, "modulus": p, "residue": int(m.group(1))})
    print(n, p, m.group(1), round(time.time() - t, 1), flush=True)
(ROOT / "tests" / "golden" / "ggcount_residues.json").write_text(json.dumps(out, indent=1) + "\n")
