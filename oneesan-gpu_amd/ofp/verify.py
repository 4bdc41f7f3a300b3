"""Verify an exact-result manifest from first principles.

    python -m ofp.verify WORK/exact_manifest.json [--binary PATH] [--verify-sources]

Checks, without trusting anything the manifest states (the checks of the
reference's scripts/tools/verify_exact_result.py:145-383, restated):

  * the manifest's own SHA-256 and its exact field set;
  * the path bound, recomputed here with a different evaluation of the same
    checkerboard-strip bound (strip counts as column-transfer matrix powers,
    partition by dynamic programming), and its bit length;
  * every congruence: a prime modulus (deterministic Miller-Rabin), no repeats,
    a canonical residue, a finite non-negative wall time;
  * the CRT: the product of the moduli exceeds the bound, the last modulus is
    needed, the product and the reconstructed integer equal the manifest's,
    the integer is within the bound and meets every congruence;
  * the wall-time sum; the solver identity's shape (and, with --binary, the
    binary's SHA-256);
  * the checkpoint file: its SHA-256, its strict format, and the same residue,
    wall time and log for every modulus used;
  * exact.txt: its SHA-256 and its exact contents;
  * the build provenance sidecar, when the manifest names one: its SHA-256 and
    strict format, its binary hash equal to the solver identity's, and with
    --verify-sources every recorded source file of this checkout.

Exit status 0 and `valid=1` on success; 1 and `INVALID: ...` otherwise.
"""
from __future__ import annotations

import argparse
import math
import sys
from pathlib import Path

from . import provenance
from .exact_safety import (
    RESULT_CHECKSUM_FIELD, file_sha256, load_checkpoint, load_exact_result_manifest,
)

MANIFEST_FIELDS = {
    "format", "n", "exact_decimal", "path_bound_decimal", "bound_bits", "bound_method",
    "strip_partition", "modulus_product_decimal", "modulus_bits", "primes_used", "congruences",
    "solver", "solver_wall_s_sum", "checkpoint_file", "checkpoint_sha256", "exact_file",
    "exact_sha256", "build_provenance_file", "build_provenance_sha256",
    "admission_certificate_file", "admission_certificate_sha256", "admission_certificate_schema",
    RESULT_CHECKSUM_FIELD,
}


# ---- the bound, evaluated independently of ofp.path_bound ----

def _strip_transfer(height: int) -> list[list[int]]:
    """0/1 matrix over column patterns of a strip `height` cells tall: columns x
    and y may stand side by side unless some row boundary is crossed by a
    vertical edge in both of them and the two columns differ in the cells
    above and below it (the checkerboard-strip rule of
    scripts/solve/path_bound.py)."""
    size = 1 << height
    low = (1 << (height - 1)) - 1 if height > 1 else 0
    t = [[1] * size for _ in range(size)]
    if height <= 1:
        return t
    for x in range(size):
        vx = (x ^ (x >> 1)) & low
        for y in range(size):
            vy = (y ^ (y >> 1)) & low
            if vx & vy & (x ^ y) & low:
                t[x][y] = 0
    return t


def _mat_vec(m: list[list[int]], v: list[int]) -> list[int]:
    return [sum(a * b for a, b in zip(row, v) if a) for row in m]


def strip_count(height: int, width: int) -> int:
    """Sum over all sequences of `width` compatible columns: 1^T T^(width-1) 1,
    evaluated as repeated matrix-vector products from the right."""
    t = _strip_transfer(height)
    v = [1] * len(t)
    for _ in range(width - 1):
        v = _mat_vec(t, v)
    return sum(v)


def path_bound(n: int, max_height: int = 9) -> tuple[int, list[int]]:
    if n < 1:
        return 1, []
    counts = {h: strip_count(h, n) for h in range(1, min(max_height, n) + 1)}
    best: dict[int, tuple[int, list[int]]] = {0: (1, [])}
    for rows in range(1, n + 1):
        cands = [(best[rows - h][0] * c, best[rows - h][1] + [h])
                 for h, c in counts.items() if h <= rows and rows - h in best]
        # the smallest product; among equals the first strip height tried (ascending)
        best[rows] = min(cands, key=lambda t: t[0])
    return best[n]


def is_prime(n: int) -> bool:
    if n < 2:
        return False
    for p in (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37):
        if n % p == 0:
            return n == p
    d, s = n - 1, 0
    while d % 2 == 0:
        d //= 2
        s += 1
    for a in (2, 325, 9375, 28178, 450775, 9780504, 1795265022):  # deterministic below 2^64
        a %= n
        if a == 0:
            continue
        x = pow(a, d, n)
        if x in (1, n - 1):
            continue
        for _ in range(s - 1):
            x = x * x % n
            if x == n - 1:
                break
        else:
            return False
    return True


def _decimal(data: dict, key: str) -> int:
    v = data.get(key)
    if not isinstance(v, str) or not v.isdigit() or str(int(v)) != v:
        raise ValueError(f"{key} must be a canonical non-negative decimal string")
    return int(v)


def _sha_field(data: dict, key: str) -> str:
    v = data[key]
    if not isinstance(v, str) or len(v) != 64 or any(c not in "0123456789abcdef" for c in v):
        raise ValueError(f"invalid {key}")
    return v


def _basename(data: dict, key: str) -> str:
    v = data[key]
    if not isinstance(v, str) or not v or Path(v).name != v:
        raise ValueError(f"{key} must be a basename")
    return v


def _wall(v, what: str) -> float:
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(float(v)) or v < 0:
        raise ValueError(f"{what}: invalid wall_s")
    return float(v)


def verify(manifest: Path, binary: Path | None = None, verify_sources: bool = False) -> dict:
    manifest = Path(manifest)
    data = load_exact_result_manifest(manifest)
    if set(data) != MANIFEST_FIELDS:
        raise ValueError(f"manifest field mismatch: missing={sorted(MANIFEST_FIELDS - set(data))} "
                         f"unknown={sorted(set(data) - MANIFEST_FIELDS)}")
    n = data["n"]
    if type(n) is not int or n < 1:
        raise ValueError("manifest n must be a positive integer")
    if data["bound_method"] != "checkerboard_strip_v1":
        raise ValueError(f"unsupported bound method {data['bound_method']!r}")
    exact = _decimal(data, "exact_decimal")
    bound = _decimal(data, "path_bound_decimal")
    product = _decimal(data, "modulus_product_decimal")
    want_bound, want_parts = path_bound(n)
    if bound != want_bound:
        raise ValueError(f"path bound mismatch: {bound} != recomputed {want_bound}")
    if data["strip_partition"] != want_parts:
        raise ValueError(f"strip partition mismatch: {data['strip_partition']} != {want_parts}")
    if data["bound_bits"] != bound.bit_length() or data["modulus_bits"] != product.bit_length():
        raise ValueError("bound_bits / modulus_bits mismatch")

    cong = data["congruences"]
    if not isinstance(cong, list) or not cong:
        raise ValueError("congruences must be a non-empty list")
    if type(data["primes_used"]) is not int or data["primes_used"] != len(cong):
        raise ValueError("primes_used mismatch")
    x, m, wall, moduli = 0, 1, 0.0, []
    for i, rec in enumerate(cong, 1):
        if not isinstance(rec, dict) or not {"modulus", "residue", "wall_s"} <= set(rec) \
                or set(rec) - {"modulus", "residue", "wall_s", "log"}:
            raise ValueError(f"congruence {i}: field mismatch")
        p, r = rec["modulus"], rec["residue"]
        if type(p) is not int or not 2 <= p < 1 << 64 or not is_prime(p):
            raise ValueError(f"congruence {i}: modulus {p!r} is not a prime below 2^64")
        if p in moduli:
            raise ValueError(f"congruence {i}: duplicate modulus {p}")
        if type(r) is not int or not 0 <= r < p:
            raise ValueError(f"congruence {i}: non-canonical residue {r!r}")
        if "log" in rec and not isinstance(rec["log"], str):
            raise ValueError(f"congruence {i}: log must be a string")
        wall += _wall(rec["wall_s"], f"congruence {i}")
        # x' = x + m * ((r - x) / m mod p): the moduli are distinct primes, so coprime
        x += m * ((r - x) * pow(m, -1, p) % p)
        m *= p
        moduli.append(p)
    if m <= bound:
        raise ValueError(f"CRT modulus product {m} does not exceed the path bound {bound}")
    if len(moduli) > 1 and m // moduli[-1] > bound:
        raise ValueError("the last congruence is not needed to exceed the path bound")
    if m != product:
        raise ValueError(f"modulus product mismatch: {product} != recomputed {m}")
    if x != exact:
        raise ValueError(f"exact value mismatch: manifest {exact} != CRT {x}")
    if exact > bound:
        raise ValueError("exact value exceeds the path bound")
    for p, rec in zip(moduli, cong):
        if exact % p != rec["residue"]:
            raise ValueError(f"exact value fails the congruence mod {p}")
    reported = _wall(data["solver_wall_s_sum"], "solver_wall_s_sum")
    if not math.isclose(reported, wall, rel_tol=1e-12, abs_tol=1e-9):
        raise ValueError(f"solver wall-time sum mismatch: {reported} != {wall}")

    solver = data["solver"]
    if not isinstance(solver, dict) or set(solver) != {"binary_name", "binary_sha256", "git_commit"} \
            or not all(isinstance(v, str) for v in solver.values()):
        raise ValueError("solver identity field mismatch")
    _sha_field(solver, "binary_sha256")
    if binary is not None and file_sha256(Path(binary)) != solver["binary_sha256"]:
        raise ValueError("solver binary SHA-256 mismatch")

    adm = [data["admission_certificate_file"], data["admission_certificate_sha256"],
           data["admission_certificate_schema"]]
    if any(v is not None for v in adm):
        raise ValueError("admission certificates are not produced by this solver")

    ck = manifest.parent / _basename(data, "checkpoint_file")
    if not ck.is_file():
        raise ValueError(f"checkpoint named by the manifest is missing: {ck}")
    if file_sha256(ck) != _sha_field(data, "checkpoint_sha256"):
        raise ValueError("checkpoint SHA-256 mismatch")
    saved = load_checkpoint(ck, n=n, identity=solver)
    if set(saved) != set(moduli):
        raise ValueError(f"checkpoint moduli {sorted(saved)} != manifest moduli {sorted(moduli)}")
    for rec in cong:
        s = saved[rec["modulus"]]
        if s.get("residue") != rec["residue"] or s.get("log") != rec.get("log") or \
                not math.isclose(float(s.get("wall_s", 0.0)), float(rec["wall_s"]), rel_tol=1e-12, abs_tol=1e-9):
            raise ValueError(f"checkpoint and manifest disagree for modulus {rec['modulus']}")

    ex = manifest.parent / _basename(data, "exact_file")
    if not ex.is_file():
        raise ValueError(f"exact file named by the manifest is missing: {ex}")
    if file_sha256(ex) != _sha_field(data, "exact_sha256"):
        raise ValueError("exact file SHA-256 mismatch")
    text = (f"n={n}\nexact={exact}\nbound_bits={bound.bit_length()}\nmodulus_bits={m.bit_length()}\n"
            f"primes_used={len(cong)}\nsolver_wall_s_sum={reported:.9f}\nmanifest_file={manifest.name}\n")
    if ex.read_text(encoding="utf-8") != text:
        raise ValueError("exact file contents do not match the verified manifest")

    pname, psha = data["build_provenance_file"], data["build_provenance_sha256"]
    if (pname is None) != (psha is None):
        raise ValueError("build provenance file and hash must both be null or both be present")
    if pname is not None:
        pf = manifest.parent / _basename(data, "build_provenance_file")
        if not pf.is_file():
            raise ValueError(f"build provenance named by the manifest is missing: {pf}")
        if file_sha256(pf) != _sha_field(data, "build_provenance_sha256"):
            raise ValueError("build provenance SHA-256 mismatch")
        pdata = provenance.load(pf, binary=binary, verify_sources=verify_sources)
        if pdata["binary_sha256"] != solver["binary_sha256"]:
            raise ValueError("build provenance binary SHA-256 differs from the solver identity")
    return data


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="verify an exact-result manifest from first principles")
    ap.add_argument("manifest", type=Path)
    ap.add_argument("--binary", type=Path, default=None, help="also check the solver binary's SHA-256")
    ap.add_argument("--verify-sources", action="store_true",
                    help="also check the build provenance's source files against this checkout")
    a = ap.parse_args(argv)
    try:
        d = verify(a.manifest, a.binary, a.verify_sources)
    except (ValueError, OSError) as exc:
        print(f"INVALID: {exc}", file=sys.stderr)
        return 1
    print(f"format={d['format']}\nn={d['n']}\nexact={d['exact_decimal']}\nprimes_used={d['primes_used']}\n"
          f"manifest_sha256={d[RESULT_CHECKSUM_FIELD]}\nvalid=1")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
