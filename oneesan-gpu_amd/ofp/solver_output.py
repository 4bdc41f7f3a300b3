"""Strict parser for the solver's machine-readable result line.

Restates the reference's contract (scripts/solve/solver_output.py:9-74,
pinned by tests/test_solver_output.py): a stdout line is a result record iff
it carries any of residue= / modulus= / wall_s=; then n, residue, modulus and
wall_s must each appear exactly once, integers in canonical decimal, wall_s a
finite non-negative float in C++ iostream syntax.  Other key=value tokens
(backend=, residue_index=, batch_wall_s=, ...) are allowed and ignored.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass

_KEYS = ("n", "residue", "modulus", "wall_s")
_MARKERS = ("residue", "modulus", "wall_s")
_UINT = re.compile(r"(?:0|[1-9][0-9]*)\Z")
_FLOAT = re.compile(r"(?:0|[1-9][0-9]*)(?:\.[0-9]+)?(?:[eE][+-]?[0-9]+)?\Z")


@dataclass(frozen=True)
class SolverResult:
    n: int
    residue: int
    modulus: int
    wall_s: float


def parse_result_line(line: str) -> SolverResult | None:
    """None for diagnostic lines; ValueError for a malformed result line."""
    values: dict[str, str] = {}
    for token in line.strip().split():
        key, sep, val = token.partition("=")
        if not sep or key not in _KEYS:
            continue
        if key in values:
            raise ValueError(f"duplicate solver result field: {key}")
        values[key] = val
    if not any(k in values for k in _MARKERS):
        return None
    missing = [k for k in _KEYS if k not in values]
    if missing:
        raise ValueError(f"incomplete solver result fields: missing={missing}")
    for k in ("n", "residue", "modulus"):
        if not _UINT.fullmatch(values[k]):
            raise ValueError(f"non-canonical solver {k}: {values[k]!r}")
    if not _FLOAT.fullmatch(values["wall_s"]):
        raise ValueError(f"invalid solver wall_s syntax: {values['wall_s']!r}")
    n, residue, modulus = int(values["n"]), int(values["residue"]), int(values["modulus"])
    wall = float(values["wall_s"])
    if n < 1:
        raise ValueError(f"invalid solver n: {n}")
    if modulus < 2:
        raise ValueError(f"invalid solver modulus: {modulus}")
    if not math.isfinite(wall) or wall < 0:
        raise ValueError("solver wall_s must be finite and nonnegative")
    return SolverResult(n=n, residue=residue, modulus=modulus, wall_s=wall)
