"""Write tests/golden/exact_counts.json: A007764(n) for n=1..12 from the
independent vertex/component DP (oracle/exact_dp.cpp), cross-checked against
the Grid-FP oracle mod two primes.  Run from the repository root:
    python tests/golden/make_exact_counts.py
"""
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import binding  # noqa: E402

out = {"_about": "A007764(n) from oracle/exact_dp.cpp (vertex-degree/component DP, unsigned __int128); "
                 "each value also equals the Grid-FP oracle (oracle/gridfp_oracle.c) mod 4294967291 and 4294966997.",
       "counts": {}}
for n in range(1, 13):
    t = time.time()
    c = binding.exact_count(n)
    for p in (4294967291, 4294966997):
        assert binding.solve(n, p) == c % p, (n, p)
    out["counts"][str(n)] = str(c)
    print(n, c, round(time.time() - t, 1), flush=True)
(ROOT / "tests" / "golden" / "exact_counts.json").write_text(json.dumps(out, indent=1) + "\n")
