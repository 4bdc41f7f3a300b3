"""Exact-count durability: residue checkpoint, result manifest, work-dir lock.

The on-disk formats are the reference's, so a run can be resumed or verified
by either implementation (reference: scripts/solve/exact_safety.py):

checkpoint.json  (format_version 4; the reference docs call it "V3",
                  README.md:36, while its code writes 4, exact_safety.py:16)
  {"format_version": 4, "n": N,
   "solver": {"binary_name", "binary_sha256", "git_commit"},
   "admission_certificate": null,
   "residues": {"<p>": {"residue": r, "wall_s": s[, "log": str]}},
   "checkpoint_sha256": sha256 of the canonical JSON of everything else}
  Strict load: duplicate keys, unknown or missing fields, non-integer or
  non-canonical values, a changed solver binary or a checksum mismatch are
  errors (exact_safety.py:146-310).  Writes are fsync + atomic rename + fsync
  of the directory (exact_safety.py:313-336).

exact_manifest.json (format ONEESAN_EXACT_RESULT_V4) and exact.txt, manifest
  first, exact.txt last as the completion marker (exact_safety.py:395-563).

Not carried over: build-provenance sidecars and row-8 admission certificates
(they certify reference build artefacts that this solver does not have); the
manifest keeps those fields, set to null.
"""
from __future__ import annotations

import fcntl
import hashlib
import hmac
import json
import math
import os
from pathlib import Path
from typing import Any, TextIO

CHECKPOINT_FORMAT_VERSION = 4
CHECKSUM_FIELD = "checkpoint_sha256"
RESULT_MANIFEST_FORMAT = "ONEESAN_EXACT_RESULT_V4"
RESULT_CHECKSUM_FIELD = "manifest_sha256"


def file_sha256(path: Path) -> str:
    h = hashlib.sha256()
    with Path(path).open("rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def solver_identity(binary: Path, git_commit: str = "unknown") -> dict[str, str]:
    """Identity a checkpoint is bound to: a changed binary invalidates resume."""
    binary = Path(binary).resolve()
    return {"binary_name": binary.name, "binary_sha256": file_sha256(binary), "git_commit": git_commit}


def validate_residue(modulus: int, residue: int, *, source: str = "residue") -> None:
    if modulus < 2:
        raise ValueError(f"{source}: invalid modulus {modulus}")
    if not 0 <= residue < modulus:
        raise ValueError(f"{source}: residue {residue} is outside canonical range [0, {modulus})")


def _validate_wall(value: Any, source: str) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise ValueError(f"{source}: wall_s must be numeric")
    w = float(value)
    if not math.isfinite(w) or w < 0:
        raise ValueError(f"{source}: wall_s must be finite and nonnegative")
    return w


def validate_residue_records(residues: dict[int, dict]) -> None:
    for modulus, rec in residues.items():
        if type(modulus) is not int:
            raise ValueError(f"checkpoint modulus key must be int, got {type(modulus).__name__}")
        if not isinstance(rec, dict):
            raise ValueError(f"checkpoint modulus {modulus}: record must be an object")
        if "residue" not in rec:
            raise ValueError(f"checkpoint modulus {modulus}: missing residue")
        if type(rec["residue"]) is not int:
            raise ValueError(f"checkpoint modulus {modulus}: residue must be a JSON integer")
        validate_residue(modulus, rec["residue"], source=f"checkpoint modulus {modulus}")
        if "wall_s" in rec:
            _validate_wall(rec["wall_s"], f"checkpoint modulus {modulus}")
        if "log" in rec and not isinstance(rec["log"], str):
            raise ValueError(f"checkpoint modulus {modulus}: log must be a string")


def _no_duplicates(pairs):
    out: dict[str, Any] = {}
    for k, v in pairs:
        if k in out:
            raise ValueError(f"duplicate JSON key in exact checkpoint: {k}")
        out[k] = v
    return out


def _canonical(data: dict[str, Any], drop: str) -> bytes:
    payload = {k: v for k, v in data.items() if k != drop}
    return json.dumps(payload, sort_keys=True, separators=(",", ":"), ensure_ascii=True).encode()


def checkpoint_checksum(data: dict[str, Any]) -> str:
    return hashlib.sha256(_canonical(data, CHECKSUM_FIELD)).hexdigest()


def result_manifest_checksum(data: dict[str, Any]) -> str:
    return hashlib.sha256(_canonical(data, RESULT_CHECKSUM_FIELD)).hexdigest()


def _load_json_strict(path: Path, what: str = "exact checkpoint") -> dict[str, Any]:
    try:
        raw = path.read_text(encoding="utf-8")
    except OSError as exc:
        raise ValueError(f"cannot read {what} {path}: {exc}") from exc

    def reject_constant(token: str):
        raise ValueError(f"non-standard JSON constant: {token}")
    try:
        data = json.loads(raw, object_pairs_hook=_no_duplicates, parse_constant=reject_constant)
    except (json.JSONDecodeError, ValueError) as exc:
        raise ValueError(f"malformed {what} {path}: {exc}") from exc
    if not isinstance(data, dict):
        raise ValueError(f"{what} {path}: root must be a JSON object")
    return data


def load_checkpoint(path: Path, *, n: int, identity: dict[str, str]) -> dict[int, dict]:
    path = Path(path)
    if not path.exists():
        return {}
    data = _load_json_strict(path)
    if type(data.get("format_version")) is not int or type(data.get("n")) is not int:
        raise ValueError(f"checkpoint {path}: format_version and n must be JSON integers")
    if data["format_version"] != CHECKPOINT_FORMAT_VERSION:
        raise ValueError(f"checkpoint {path}: unsupported format_version {data['format_version']}")
    saved = data.get(CHECKSUM_FIELD)
    if not isinstance(saved, str) or len(saved) != 64:
        raise ValueError(f"checkpoint {path} has no integrity checksum")
    if not hmac.compare_digest(saved, checkpoint_checksum(data)):
        raise ValueError(f"checkpoint {path}: SHA-256 integrity checksum mismatch")
    if data["n"] != n:
        raise ValueError(f"checkpoint {path} belongs to n={data['n']}, expected n={n}")
    expected = {"format_version", "n", "solver", "admission_certificate", "residues", CHECKSUM_FIELD}
    if set(data) - expected:
        raise ValueError(f"checkpoint {path}: unknown fields: {sorted(set(data) - expected)}")
    if expected - set(data):
        raise ValueError(f"checkpoint {path}: missing fields: {sorted(expected - set(data))}")
    solver = data["solver"]
    keys = {"binary_name", "binary_sha256", "git_commit"}
    if not isinstance(solver, dict) or set(solver) != keys or not all(isinstance(solver[k], str) for k in keys):
        raise ValueError(f"checkpoint {path}: solver identity fields mismatch")
    sha = solver["binary_sha256"]
    if len(sha) != 64 or any(c not in "0123456789abcdef" for c in sha):
        raise ValueError(f"checkpoint {path}: invalid solver binary SHA-256")
    if sha != identity["binary_sha256"]:
        raise ValueError(f"checkpoint {path}: solver binary changed ({sha!r} != {identity['binary_sha256']!r})")
    if data["admission_certificate"] is not None:
        raise ValueError(f"checkpoint {path}: admission certificates are not supported by this solver")
    raw = data["residues"]
    if not isinstance(raw, dict):
        raise ValueError(f"checkpoint {path}: residues must be an object")
    out: dict[int, dict] = {}
    for key, rec in raw.items():
        try:
            modulus = int(key)
        except ValueError as exc:
            raise ValueError(f"checkpoint {path}: invalid modulus key {key!r}") from exc
        if str(modulus) != key:
            raise ValueError(f"checkpoint {path}: non-canonical modulus key {key!r}")
        if not isinstance(rec, dict):
            raise ValueError(f"checkpoint modulus {modulus}: record must be an object")
        unknown = sorted(set(rec) - {"residue", "wall_s", "log"})
        if unknown:
            raise ValueError(f"checkpoint modulus {modulus}: unknown fields: {unknown}")
        out[modulus] = rec
    validate_residue_records(out)
    return out


def durable_atomic_write(path: Path, text: str) -> None:
    path = Path(path)
    tmp = path.with_suffix(path.suffix + ".tmp")
    fd = os.open(tmp, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    try:
        with os.fdopen(fd, "w", encoding="utf-8") as f:
            fd = -1
            f.write(text)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
        dfd = os.open(path.parent, os.O_RDONLY | os.O_DIRECTORY)
        try:
            os.fsync(dfd)
        finally:
            os.close(dfd)
    except Exception:
        if fd >= 0:
            os.close(fd)
        try:
            tmp.unlink()
        except FileNotFoundError:
            pass
        raise


def save_checkpoint(path: Path, *, n: int, identity: dict[str, str], residues: dict[int, dict]) -> None:
    validate_residue_records(residues)
    payload: dict[str, Any] = {
        "format_version": CHECKPOINT_FORMAT_VERSION,
        "n": n,
        "solver": identity,
        "admission_certificate": None,
        "residues": {str(k): v for k, v in residues.items()},
    }
    payload[CHECKSUM_FIELD] = checkpoint_checksum(payload)
    durable_atomic_write(path, json.dumps(payload, indent=2, sort_keys=True, ensure_ascii=True) + "\n")


def crt_pair(x: int, m: int, r: int, p: int) -> tuple[int, int]:
    """Combine x mod m with r mod p (coprime) into the residue mod m*p."""
    if m == 1:
        return r % p, p
    t = ((r - x) * pow(m, -1, p)) % p
    return x + m * t, m * p


def validate_exact_reconstruction(x: int, modulus_product: int, path_bound: int,
                                  used: list[tuple[int, int]]) -> None:
    if modulus_product <= path_bound:
        raise ValueError(f"CRT capacity insufficient: M={modulus_product} <= path_bound={path_bound}")
    if not 0 <= x <= path_bound:
        raise ValueError(f"CRT reconstruction violates rigorous path bound: exact={x}, bound={path_bound}")
    for p, r in used:
        validate_residue(p, r, source=f"final modulus {p}")
        if x % p != r:
            raise ValueError(f"CRT reconstruction failed final congruence check for modulus {p}")


def acquire_workdir_lock(work: Path) -> TextIO:
    fd = (Path(work) / ".solve.lock").open("a+")
    try:
        fcntl.flock(fd.fileno(), fcntl.LOCK_EX | fcntl.LOCK_NB)
    except BlockingIOError as exc:
        fd.close()
        raise RuntimeError(f"work directory is already in use: {work}") from exc
    return fd


def write_exact_result(work: Path, *, n: int, exact: int, path_bound: int, modulus_product: int,
                       used_moduli: list[int], residues: dict[int, dict], identity: dict[str, str],
                       checkpoint_path: Path, total_wall_s: float,
                       strip_partition: list[int],
                       build_provenance: Path | None = None) -> tuple[Path, Path]:
    validate_residue_records(residues)
    if len(set(used_moduli)) != len(used_moduli):
        raise ValueError("exact result contains duplicate CRT moduli")
    missing = [p for p in used_moduli if p not in residues]
    if missing:
        raise ValueError(f"exact result is missing residues for moduli: {missing}")
    prod = 1
    for p in used_moduli:
        prod *= p
    if prod != modulus_product:
        raise ValueError("CRT modulus product mismatch before finalization")
    validate_exact_reconstruction(exact, modulus_product, path_bound,
                                  [(p, residues[p]["residue"]) for p in used_moduli])
    if not Path(checkpoint_path).exists():
        raise ValueError(f"cannot finalize exact result without checkpoint: {checkpoint_path}")
    congruences = []
    for p in used_moduli:
        rec = residues[p]
        item: dict[str, Any] = {"modulus": p, "residue": rec["residue"], "wall_s": float(rec.get("wall_s", 0.0))}
        if "log" in rec:
            item["log"] = rec["log"]
        congruences.append(item)
    wall_sum = sum(c["wall_s"] for c in congruences)
    if not math.isclose(float(total_wall_s), wall_sum, rel_tol=1e-12, abs_tol=1e-9):
        raise ValueError(f"solver wall-time sum mismatch before finalization: {total_wall_s} != {wall_sum}")
    work = Path(work)
    manifest_path, exact_path = work / "exact_manifest.json", work / "exact.txt"
    exact_text = (f"n={n}\nexact={exact}\nbound_bits={path_bound.bit_length()}\n"
                  f"modulus_bits={modulus_product.bit_length()}\nprimes_used={len(used_moduli)}\n"
                  f"solver_wall_s_sum={total_wall_s:.9f}\nmanifest_file={manifest_path.name}\n")
    payload: dict[str, Any] = {
        "format": RESULT_MANIFEST_FORMAT,
        "n": n,
        "exact_decimal": str(exact),
        "path_bound_decimal": str(path_bound),
        "bound_bits": path_bound.bit_length(),
        "bound_method": "checkerboard_strip_v1",
        "strip_partition": strip_partition,
        "modulus_product_decimal": str(modulus_product),
        "modulus_bits": modulus_product.bit_length(),
        "primes_used": len(used_moduli),
        "congruences": congruences,
        "solver": identity,
        "solver_wall_s_sum": float(total_wall_s),
        "checkpoint_file": Path(checkpoint_path).name,
        "checkpoint_sha256": file_sha256(checkpoint_path),
        "exact_file": exact_path.name,
        "exact_sha256": hashlib.sha256(exact_text.encode()).hexdigest(),
        "build_provenance_file": Path(build_provenance).name if build_provenance else None,
        "build_provenance_sha256": file_sha256(build_provenance) if build_provenance else None,
        "admission_certificate_file": None,
        "admission_certificate_sha256": None,
        "admission_certificate_schema": None,
    }
    payload[RESULT_CHECKSUM_FIELD] = result_manifest_checksum(payload)
    durable_atomic_write(manifest_path, json.dumps(payload, indent=2, sort_keys=True, ensure_ascii=True) + "\n")
    durable_atomic_write(exact_path, exact_text)
    return exact_path, manifest_path


def load_exact_result_manifest(path: Path) -> dict[str, Any]:
    data = _load_json_strict(Path(path), "exact result manifest")
    saved = data.get(RESULT_CHECKSUM_FIELD)
    if not isinstance(saved, str) or len(saved) != 64:
        raise ValueError(f"exact result manifest {path}: missing SHA-256 checksum")
    if not hmac.compare_digest(saved, result_manifest_checksum(data)):
        raise ValueError(f"exact result manifest {path}: SHA-256 integrity checksum mismatch")
    if data.get("format") != RESULT_MANIFEST_FORMAT:
        raise ValueError(f"exact result manifest {path}: unsupported format {data.get('format')!r}")
    return data
