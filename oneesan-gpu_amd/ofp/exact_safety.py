"""Durable state of an exact (CRT) run: the residue checkpoint, the result
manifest and exact.txt, and the work-directory lock.

The files are interchangeable with the reference's (scripts/solve/exact_safety.py),
so either implementation can resume or verify a run:

checkpoint.json -- format_version 4 (the reference's docs call the format "V3",
  README.md:36; its code writes 4, exact_safety.py:16):

      format_version  int, 4
      n               int
      solver          {binary_name: str, binary_sha256: 64 hex, git_commit: str}
      admission_certificate   null (this solver has no row-8 automaton to certify)
      residues        {"<modulus>": {residue: int in [0, modulus), wall_s?: finite >= 0, log?: str}}
      checkpoint_sha256       SHA-256 of the canonical JSON of all other fields

exact_manifest.json -- format ONEESAN_EXACT_RESULT_V4 with manifest_sha256 over
  the canonical JSON of its other fields; exact.txt -- the human-readable
  result, written after the manifest, so that its presence marks completion.

Canonical JSON = sorted keys, "," and ":" separators, ASCII.  Every file is
replaced atomically: written to a temporary file in the same directory, flushed
to disk, renamed over the target, and the directory entry flushed too.

Reading is strict: a duplicate key, a NaN/Infinity constant, a missing or
unknown field, a value of the wrong JSON type, a non-canonical modulus key, a
residue outside [0, modulus), a changed solver binary or a checksum mismatch
is an error.  The phrases the reference's tests match on
(tests/test_exact_safety.py, tests/test_exact_manifest.py) are kept in the
messages: "duplicate JSON key", "malformed exact checkpoint", "format_version
and n must be JSON integers", "residue must be a JSON integer", "outside
canonical range", "SHA-256 integrity checksum mismatch", "solver binary
changed", "violates rigorous path bound", "final congruence check", "already in
use".
"""
from __future__ import annotations

import fcntl
import hashlib
import hmac
import json
import math
import os
import tempfile
from pathlib import Path
from typing import Any, Callable, TextIO

CHECKPOINT_FORMAT_VERSION = 4
CHECKSUM_FIELD = "checkpoint_sha256"
RESULT_MANIFEST_FORMAT = "ONEESAN_EXACT_RESULT_V4"
RESULT_CHECKSUM_FIELD = "manifest_sha256"
LOCK_NAME = ".solve.lock"

_HEX = frozenset("0123456789abcdef")


# ---------------------------------------------------------------- hashing

def file_sha256(path: Path) -> str:
    digest = hashlib.sha256()
    with open(path, "rb") as f:
        while block := f.read(1 << 20):
            digest.update(block)
    return digest.hexdigest()


def canonical_json(obj: Any) -> bytes:
    return json.dumps(obj, sort_keys=True, separators=(",", ":"), ensure_ascii=True).encode("ascii")


def _sha_without(data: dict[str, Any], field: str) -> str:
    return hashlib.sha256(canonical_json({k: v for k, v in data.items() if k != field})).hexdigest()


def checkpoint_checksum(data: dict[str, Any]) -> str:
    return _sha_without(data, CHECKSUM_FIELD)


def result_manifest_checksum(data: dict[str, Any]) -> str:
    return _sha_without(data, RESULT_CHECKSUM_FIELD)


def _is_sha256(value: Any) -> bool:
    return isinstance(value, str) and len(value) == 64 and set(value) <= _HEX


def solver_identity(binary: Path, git_commit: str = "unknown") -> dict[str, str]:
    """What a checkpoint is bound to: resuming with a different binary is refused."""
    resolved = Path(binary).resolve()
    return {"binary_name": resolved.name, "binary_sha256": file_sha256(resolved), "git_commit": git_commit}


# ---------------------------------------------------------------- value checks

def _is_int(v: Any) -> bool:
    return type(v) is int  # JSON integers only: bool is not one


def validate_residue(modulus: int, residue: int, *, source: str = "residue") -> None:
    if not _is_int(modulus) or modulus < 2:
        raise ValueError(f"{source}: modulus {modulus!r} is not an integer >= 2")
    if not _is_int(residue) or not 0 <= residue < modulus:
        raise ValueError(f"{source}: residue {residue!r} is outside canonical range [0, {modulus})")


def _wall_seconds(value: Any, source: str) -> float:
    if isinstance(value, bool) or not isinstance(value, (int, float)):
        raise ValueError(f"{source}: wall_s is not a number")
    seconds = float(value)
    if not (math.isfinite(seconds) and seconds >= 0.0):
        raise ValueError(f"{source}: wall_s {value!r} is not a finite, non-negative number")
    return seconds


RECORD_FIELDS = {"residue", "wall_s", "log"}


def validate_residue_records(residues: dict[int, dict]) -> None:
    """Every record: an int modulus key, a JSON-integer residue in range, optional
    finite non-negative wall_s and string log, nothing else."""
    for modulus, rec in residues.items():
        where = f"checkpoint modulus {modulus}"
        if not _is_int(modulus):
            raise ValueError(f"checkpoint modulus key {modulus!r} is not an int")
        if not isinstance(rec, dict):
            raise ValueError(f"{where}: record is not an object")
        extra = set(rec) - RECORD_FIELDS
        if extra:
            raise ValueError(f"{where}: unexpected fields {sorted(extra)}")
        if "residue" not in rec:
            raise ValueError(f"{where}: no residue")
        if not _is_int(rec["residue"]):
            raise ValueError(f"{where}: residue must be a JSON integer")
        validate_residue(modulus, rec["residue"], source=where)
        if "wall_s" in rec:
            _wall_seconds(rec["wall_s"], where)
        if "log" in rec and not isinstance(rec["log"], str):
            raise ValueError(f"{where}: log is not a string")


# ---------------------------------------------------------------- strict JSON

def _pairs_without_duplicates(pairs: list[tuple[str, Any]]) -> dict[str, Any]:
    seen: dict[str, Any] = {}
    for key, value in pairs:
        if key in seen:
            raise ValueError(f"duplicate JSON key {key!r}")
        seen[key] = value
    return seen


def _refuse_constant(name: str) -> Any:
    raise ValueError(f"JSON constant {name} is not allowed")


def read_strict_json(path: Path, what: str) -> dict[str, Any]:
    """A JSON object from `path`; any defect is a ValueError naming `what`."""
    try:
        text = Path(path).read_text(encoding="utf-8")
    except OSError as exc:
        raise ValueError(f"{what} {path} cannot be read: {exc}") from exc
    try:
        obj = json.loads(text, object_pairs_hook=_pairs_without_duplicates, parse_constant=_refuse_constant)
    except ValueError as exc:  # JSONDecodeError is a ValueError
        raise ValueError(f"malformed {what} {path}: {exc}") from exc
    if not isinstance(obj, dict):
        raise ValueError(f"malformed {what} {path}: the top level is not an object")
    return obj


Check = Callable[[Any], bool]


def _require_fields(obj: dict[str, Any], spec: dict[str, Check], where: str) -> None:
    """obj has exactly the keys of spec, each value passing its check."""
    missing = sorted(set(spec) - set(obj))
    if missing:
        raise ValueError(f"{where}: fields missing: {missing}")
    unknown = sorted(set(obj) - set(spec))
    if unknown:
        raise ValueError(f"{where}: fields not in the format: {unknown}")
    for key, ok in spec.items():
        if not ok(obj[key]):
            raise ValueError(f"{where}: field {key!r} has an invalid value {obj[key]!r}")


SOLVER_SPEC: dict[str, Check] = {
    "binary_name": lambda v: isinstance(v, str),
    "binary_sha256": _is_sha256,
    "git_commit": lambda v: isinstance(v, str),
}
CHECKPOINT_SPEC: dict[str, Check] = {
    "format_version": _is_int,
    "n": _is_int,
    "solver": lambda v: isinstance(v, dict),
    "admission_certificate": lambda v: v is None,
    "residues": lambda v: isinstance(v, dict),
    CHECKSUM_FIELD: _is_sha256,
}


# ---------------------------------------------------------------- checkpoint

def load_checkpoint(path: Path, *, n: int, identity: dict[str, str]) -> dict[int, dict]:
    """Residues of a checkpoint ({} when there is none), after checking its
    format, checksum, n and solver identity."""
    path = Path(path)
    if not path.exists():
        return {}
    data = read_strict_json(path, "exact checkpoint")
    where = f"checkpoint {path}"
    if not (_is_int(data.get("format_version")) and _is_int(data.get("n"))):
        raise ValueError(f"{where}: format_version and n must be JSON integers")
    if data["format_version"] != CHECKPOINT_FORMAT_VERSION:
        raise ValueError(f"{where}: format_version {data['format_version']} is not "
                         f"{CHECKPOINT_FORMAT_VERSION}")
    if not _is_sha256(data.get(CHECKSUM_FIELD)):
        raise ValueError(f"{where}: no {CHECKSUM_FIELD}")
    if not hmac.compare_digest(data[CHECKSUM_FIELD], checkpoint_checksum(data)):
        raise ValueError(f"{where}: SHA-256 integrity checksum mismatch")
    _require_fields(data, CHECKPOINT_SPEC, where)
    if data["n"] != n:
        raise ValueError(f"{where} belongs to n={data['n']}, this run is n={n}")
    _require_fields(data["solver"], SOLVER_SPEC, f"{where} solver")
    saved = data["solver"]["binary_sha256"]
    if saved != identity["binary_sha256"]:
        raise ValueError(f"{where}: solver binary changed since the checkpoint was written "
                         f"({saved} -> {identity['binary_sha256']})")
    out: dict[int, dict] = {}
    for key, rec in data["residues"].items():
        if not (key.isdigit() and str(int(key)) == key):
            raise ValueError(f"{where}: modulus key {key!r} is not a canonical decimal")
        out[int(key)] = rec
    validate_residue_records(out)
    return out


def durable_atomic_write(path: Path, text: str) -> None:
    """Replace `path` with `text` so that a crash leaves the old or the new file,
    never a torn one: temporary file in the same directory, fsync, rename, fsync
    of the directory."""
    target = Path(path)
    directory = target.parent
    handle, tmp_name = tempfile.mkstemp(prefix=f".{target.name}.", suffix=".tmp", dir=directory)
    tmp = Path(tmp_name)
    try:
        with os.fdopen(handle, "wb") as out:
            out.write(text.encode("utf-8"))
            out.flush()
            os.fsync(out.fileno())
        os.chmod(tmp, 0o644)
        os.replace(tmp, target)
    except BaseException:
        tmp.unlink(missing_ok=True)
        raise
    dir_fd = os.open(directory, os.O_RDONLY | os.O_DIRECTORY)
    try:
        os.fsync(dir_fd)
    finally:
        os.close(dir_fd)


def _pretty(obj: dict[str, Any]) -> str:
    return json.dumps(obj, indent=2, sort_keys=True, ensure_ascii=True) + "\n"


def save_checkpoint(path: Path, *, n: int, identity: dict[str, str], residues: dict[int, dict]) -> None:
    validate_residue_records(residues)
    doc: dict[str, Any] = {
        "format_version": CHECKPOINT_FORMAT_VERSION,
        "n": n,
        "solver": dict(identity),
        "admission_certificate": None,
        "residues": {str(m): residues[m] for m in sorted(residues)},
    }
    doc[CHECKSUM_FIELD] = checkpoint_checksum(doc)
    durable_atomic_write(path, _pretty(doc))


# ---------------------------------------------------------------- CRT

def crt_pair(x: int, m: int, r: int, p: int) -> tuple[int, int]:
    """The residue mod m*p congruent to x mod m and r mod p (gcd(m, p) = 1)."""
    if m == 1:
        return r % p, p
    lift = (r - x) * pow(m, -1, p) % p
    return x + m * lift, m * p


def validate_exact_reconstruction(x: int, modulus_product: int, path_bound: int,
                                  used: list[tuple[int, int]]) -> None:
    """x is the count only if the moduli's product exceeds the bound, x lies in
    [0, bound] and x meets every congruence it was built from."""
    if modulus_product <= path_bound:
        raise ValueError(f"moduli product ({modulus_product.bit_length()} bits) does not exceed the path "
                         f"bound ({path_bound.bit_length()} bits)")
    if not 0 <= x <= path_bound:
        raise ValueError(f"CRT value {x} violates rigorous path bound {path_bound}")
    for p, r in used:
        validate_residue(p, r, source=f"final modulus {p}")
        if x % p != r:
            raise ValueError(f"CRT value fails the final congruence check mod {p}")


def acquire_workdir_lock(work: Path) -> TextIO:
    """An exclusive, non-blocking lock on <work>/.solve.lock, held while the
    returned file stays open."""
    handle = open(Path(work) / LOCK_NAME, "a+")
    try:
        fcntl.flock(handle, fcntl.LOCK_EX | fcntl.LOCK_NB)
    except OSError as exc:
        handle.close()
        raise RuntimeError(f"work directory {work} is already in use by another runner") from exc
    return handle


# ---------------------------------------------------------------- result

def _product(values: list[int]) -> int:
    out = 1
    for v in values:
        out *= v
    return out


def exact_text(n: int, exact: int, path_bound: int, modulus_product: int, primes: int, wall: float,
               manifest_name: str) -> str:
    """exact.txt: one key=value per line."""
    lines = [f"n={n}", f"exact={exact}", f"bound_bits={path_bound.bit_length()}",
             f"modulus_bits={modulus_product.bit_length()}", f"primes_used={primes}",
             f"solver_wall_s_sum={wall:.9f}", f"manifest_file={manifest_name}"]
    return "\n".join(lines) + "\n"


def write_exact_result(work: Path, *, n: int, exact: int, path_bound: int, modulus_product: int,
                       used_moduli: list[int], residues: dict[int, dict], identity: dict[str, str],
                       checkpoint_path: Path, total_wall_s: float,
                       strip_partition: list[int],
                       build_provenance: Path | None = None) -> tuple[Path, Path]:
    """Write exact_manifest.json, then exact.txt; returns their paths."""
    validate_residue_records(residues)
    if len(set(used_moduli)) != len(used_moduli):
        raise ValueError("a CRT modulus is used twice")
    absent = [p for p in used_moduli if p not in residues]
    if absent:
        raise ValueError(f"no residue recorded for moduli {absent}")
    if _product(used_moduli) != modulus_product:
        raise ValueError("the product of the used moduli is not the stated modulus product")
    validate_exact_reconstruction(exact, modulus_product, path_bound,
                                  [(p, residues[p]["residue"]) for p in used_moduli])
    checkpoint_path = Path(checkpoint_path)
    if not checkpoint_path.exists():
        raise ValueError(f"checkpoint {checkpoint_path} is missing: a result needs its checkpoint")
    congruences: list[dict[str, Any]] = []
    for p in used_moduli:
        entry: dict[str, Any] = {"modulus": p, "residue": residues[p]["residue"],
                                 "wall_s": float(residues[p].get("wall_s", 0.0))}
        if "log" in residues[p]:
            entry["log"] = residues[p]["log"]
        congruences.append(entry)
    wall = math.fsum(c["wall_s"] for c in congruences)
    if not math.isclose(float(total_wall_s), wall, rel_tol=1e-12, abs_tol=1e-9):
        raise ValueError(f"total solver wall time {total_wall_s} is not the sum of the residues' {wall}")
    work = Path(work)
    manifest_path, exact_path = work / "exact_manifest.json", work / "exact.txt"
    text = exact_text(n, exact, path_bound, modulus_product, len(used_moduli), float(total_wall_s),
                      manifest_path.name)
    prov = Path(build_provenance) if build_provenance else None
    manifest: dict[str, Any] = {
        "format": RESULT_MANIFEST_FORMAT,
        "n": n,
        "exact_decimal": str(exact),
        "path_bound_decimal": str(path_bound),
        "bound_bits": path_bound.bit_length(),
        "bound_method": "checkerboard_strip_v1",
        "strip_partition": list(strip_partition),
        "modulus_product_decimal": str(modulus_product),
        "modulus_bits": modulus_product.bit_length(),
        "primes_used": len(used_moduli),
        "congruences": congruences,
        "solver": dict(identity),
        "solver_wall_s_sum": float(total_wall_s),
        "checkpoint_file": checkpoint_path.name,
        "checkpoint_sha256": file_sha256(checkpoint_path),
        "exact_file": exact_path.name,
        "exact_sha256": hashlib.sha256(text.encode("utf-8")).hexdigest(),
        "build_provenance_file": prov.name if prov else None,
        "build_provenance_sha256": file_sha256(prov) if prov else None,
        "admission_certificate_file": None,
        "admission_certificate_sha256": None,
        "admission_certificate_schema": None,
    }
    manifest[RESULT_CHECKSUM_FIELD] = result_manifest_checksum(manifest)
    durable_atomic_write(manifest_path, _pretty(manifest))
    durable_atomic_write(exact_path, text)
    return exact_path, manifest_path


def load_exact_result_manifest(path: Path) -> dict[str, Any]:
    """A result manifest whose format and checksum check out."""
    data = read_strict_json(Path(path), "exact result manifest")
    where = f"exact result manifest {path}"
    if not _is_sha256(data.get(RESULT_CHECKSUM_FIELD)):
        raise ValueError(f"{where}: no {RESULT_CHECKSUM_FIELD}")
    if not hmac.compare_digest(data[RESULT_CHECKSUM_FIELD], result_manifest_checksum(data)):
        raise ValueError(f"{where}: manifest SHA-256 checksum mismatch")
    if data.get("format") != RESULT_MANIFEST_FORMAT:
        raise ValueError(f"{where}: format {data.get('format')!r} is not {RESULT_MANIFEST_FORMAT}")
    return data
