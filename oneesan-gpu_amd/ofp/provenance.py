"""Build provenance sidecar for the solver binary.

    python -m ofp.provenance create [--binary PATH]
    python -m ofp.provenance verify PROVENANCE [--binary PATH] [--verify-sources]

A sidecar records what a solver binary was built from, so an exact result can
name the build that produced it: the binary's SHA-256 and size, the CLI source
and the closure of its quoted includes, the library and its sources as
auxiliary dependencies (role, path, SHA-256, size), the compiler and its
version, the compile arguments and the git commit; a SHA-256 over the
canonical JSON of all of that guards the file.  The layout is the reference's
ONEESAN_BUILD_PROVENANCE_V2 (scripts/solve/build_provenance.py:15-18,141-190),
so the reference's loader reads it; this solver is one binary for all
n, so it carries no -DTARGET_W argument (the reference's verifier asks for one,
scripts/tools/verify_exact_result.py:358).
"""
from __future__ import annotations

import argparse
import hashlib
import hmac
import json
import re
import subprocess
import sys
from pathlib import Path
from typing import Any

from .exact_safety import durable_atomic_write, file_sha256

FORMAT_V1 = "ONEESAN_BUILD_PROVENANCE_V1"
FORMAT_V2 = "ONEESAN_BUILD_PROVENANCE_V2"
CHECKSUM_FIELD = "provenance_sha256"
SIDECAR_SUFFIX = ".provenance.json"
_INCLUDE = re.compile(r'^\s*#\s*include\s+"([^"]+)"')
_ROLE = re.compile(r"[a-z0-9][a-z0-9_.-]*")

PKG = Path(__file__).resolve().parents[1]
ROOT = PKG.parent


def include_closure(source: Path, root: Path) -> list[Path]:
    """The source and every file it reaches through quoted #include lines, all
    inside root, sorted."""
    root = root.resolve()
    todo, seen = [source.resolve()], set()
    while todo:
        f = todo.pop()
        if f in seen:
            continue
        if not f.is_file():
            raise ValueError(f"source dependency is missing: {f}")
        try:
            f.relative_to(root)
        except ValueError as exc:
            raise ValueError(f"source dependency outside the repository: {f}") from exc
        seen.add(f)
        for line in f.read_text(encoding="utf-8").splitlines():
            m = _INCLUDE.match(line)
            if m:
                todo.append((f.parent / m.group(1)).resolve())
    return sorted(seen)


def _record(path: Path, root: Path, role: str | None = None) -> dict[str, Any]:
    path = path.resolve()
    if not path.is_file():
        raise ValueError(f"provenance dependency is missing: {path}")
    rec: dict[str, Any] = {"path": path.relative_to(root.resolve()).as_posix(),
                           "sha256": file_sha256(path), "size": path.stat().st_size}
    if role is not None:
        if not _ROLE.fullmatch(role):
            raise ValueError(f"invalid dependency role {role!r}")
        rec["role"] = role
    return rec


def checksum(data: dict[str, Any]) -> str:
    body = {k: v for k, v in data.items() if k != CHECKSUM_FIELD}
    return hashlib.sha256(json.dumps(body, sort_keys=True, separators=(",", ":"),
                                     ensure_ascii=True).encode()).hexdigest()


def _compiler_version(compiler: Path) -> str:
    proc = subprocess.run([str(compiler), "--version"], text=True, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, check=False)
    text = proc.stdout.strip()
    if proc.returncode != 0 or not text:
        raise ValueError(f"{compiler} --version failed (rc={proc.returncode})")
    return text


def _git_commit(root: Path) -> str:
    proc = subprocess.run(["git", "rev-parse", "HEAD"], cwd=root, text=True, stdout=subprocess.PIPE,
                          stderr=subprocess.DEVNULL, check=False)
    return proc.stdout.strip() if proc.returncode == 0 and proc.stdout.strip() else "unknown"


def create(binary: Path, *, source: Path, compiler: Path, compile_args: list[str],
           auxiliary: list[tuple[str, Path]], root: Path = ROOT, output: Path | None = None) -> Path:
    root = root.resolve()
    binary = Path(binary).resolve()
    if not binary.is_file():
        raise ValueError(f"binary is missing: {binary}")
    aux = sorted((_record(p, root, role) for role, p in auxiliary), key=lambda r: (r["role"], r["path"]))
    if len({r["role"] for r in aux}) != len(aux) or len({r["path"] for r in aux}) != len(aux):
        raise ValueError("auxiliary dependency roles and paths must be unique")
    data: dict[str, Any] = {
        "format": FORMAT_V2,
        "binary_file": binary.name,
        "binary_sha256": file_sha256(binary),
        "binary_size": binary.stat().st_size,
        "source": Path(source).resolve().relative_to(root).as_posix(),
        "dependencies": [_record(p, root) for p in include_closure(Path(source), root)],
        "auxiliary_dependencies": aux,
        "compiler_path": str(Path(compiler).resolve()),
        "compiler_version": _compiler_version(Path(compiler)),
        "compile_args": list(compile_args),
        "git_commit": _git_commit(root),
    }
    data[CHECKSUM_FIELD] = checksum(data)
    out = Path(output) if output else Path(str(binary) + SIDECAR_SUFFIX)
    durable_atomic_write(out, json.dumps(data, indent=2, sort_keys=True, ensure_ascii=True) + "\n")
    return out


def _sha(value: Any, what: str) -> str:
    if not isinstance(value, str) or not re.fullmatch(r"[0-9a-f]{64}", value):
        raise ValueError(f"invalid {what} SHA-256")
    return value


def _check_records(recs: Any, auxiliary: bool) -> None:
    what = "auxiliary dependencies" if auxiliary else "dependencies"
    if not isinstance(recs, list):
        raise ValueError(f"build provenance {what} must be a list")
    fields = {"path", "sha256", "size", "role"} if auxiliary else {"path", "sha256", "size"}
    for i, r in enumerate(recs, 1):
        if not isinstance(r, dict) or set(r) != fields:
            raise ValueError(f"build provenance {what} {i}: field mismatch")
        p = r["path"]
        if not isinstance(p, str) or Path(p).is_absolute() or ".." in Path(p).parts:
            raise ValueError(f"build provenance {what} {i}: unsafe path")
        _sha(r["sha256"], f"{what} {i}")
        if type(r["size"]) is not int or r["size"] < 0:
            raise ValueError(f"build provenance {what} {i}: invalid size")
        if auxiliary and (not isinstance(r["role"], str) or not _ROLE.fullmatch(r["role"])):
            raise ValueError(f"build provenance {what} {i}: invalid role")
    paths = [r["path"] for r in recs]
    if len(set(paths)) != len(paths):
        raise ValueError(f"build provenance {what}: duplicate paths")
    if auxiliary:
        if len({r["role"] for r in recs}) != len(recs):
            raise ValueError("build provenance auxiliary dependencies: duplicate roles")
        if recs != sorted(recs, key=lambda r: (r["role"], r["path"])):
            raise ValueError("build provenance auxiliary dependencies must be sorted by role and path")
    elif paths != sorted(paths):
        raise ValueError("build provenance dependencies must be sorted")


def load(path: Path, *, binary: Path | None = None, verify_sources: bool = False,
         root: Path = ROOT) -> dict[str, Any]:
    """Strict load: exact field set, checksum, record shapes; with binary, its
    name, size and SHA-256; with verify_sources, every recorded file of this
    checkout (and the include closure of the source) must match."""
    def no_dupes(pairs):
        out = {}
        for k, v in pairs:
            if k in out:
                raise ValueError(f"duplicate build-provenance key {k}")
            out[k] = v
        return out
    try:
        data = json.loads(Path(path).read_text(encoding="utf-8"), object_pairs_hook=no_dupes,
                          parse_constant=lambda t: (_ for _ in ()).throw(ValueError(f"non-standard constant {t}")))
    except (OSError, ValueError) as exc:
        raise ValueError(f"malformed build provenance {path}: {exc}") from exc
    if not isinstance(data, dict):
        raise ValueError("build provenance must be a JSON object")
    common = {"format", "binary_file", "binary_sha256", "binary_size", "source", "dependencies",
              "compiler_path", "compiler_version", "compile_args", "git_commit", CHECKSUM_FIELD}
    fmt = data.get("format")
    if fmt == FORMAT_V1:
        fields = common
    elif fmt == FORMAT_V2:
        fields = common | {"auxiliary_dependencies"}
    else:
        raise ValueError(f"unsupported build provenance format {fmt!r}")
    if set(data) != fields:
        raise ValueError(f"build provenance field mismatch: missing={sorted(fields - set(data))} "
                         f"unknown={sorted(set(data) - fields)}")
    if not hmac.compare_digest(_sha(data[CHECKSUM_FIELD], "build provenance checksum"), checksum(data)):
        raise ValueError("build provenance SHA-256 checksum mismatch")
    for key in ("binary_file", "source", "compiler_path", "compiler_version", "git_commit"):
        if not isinstance(data[key], str):
            raise ValueError(f"build provenance {key} must be a string")
    if Path(data["binary_file"]).name != data["binary_file"]:
        raise ValueError("build provenance binary_file must be a basename")
    _sha(data["binary_sha256"], "build provenance binary")
    if type(data["binary_size"]) is not int or data["binary_size"] < 0:
        raise ValueError("invalid build provenance binary_size")
    if not isinstance(data["compile_args"], list) or not all(isinstance(a, str) for a in data["compile_args"]):
        raise ValueError("build provenance compile_args must be a list of strings")
    _check_records(data["dependencies"], auxiliary=False)
    _check_records(data.get("auxiliary_dependencies", []), auxiliary=True)
    if binary is not None:
        b = Path(binary).resolve()
        if b.name != data["binary_file"]:
            raise ValueError(f"provenance binary name mismatch: {b.name} != {data['binary_file']}")
        if b.stat().st_size != data["binary_size"] or file_sha256(b) != data["binary_sha256"]:
            raise ValueError("build provenance binary hash/size mismatch")
    if verify_sources:
        root = root.resolve()
        closure = [p.relative_to(root).as_posix() for p in include_closure(root / data["source"], root)]
        if closure != [r["path"] for r in data["dependencies"]]:
            raise ValueError("build provenance dependency closure mismatch")
        for r in data["dependencies"] + data.get("auxiliary_dependencies", []):
            f = (root / r["path"]).resolve()
            if not f.is_file():
                raise ValueError(f"build provenance dependency is missing: {r['path']}")
            if f.stat().st_size != r["size"] or file_sha256(f) != r["sha256"]:
                raise ValueError(f"build provenance dependency hash/size mismatch: {r['path']}")
    return data


def _build_recipe():
    import importlib.util
    spec = importlib.util.spec_from_file_location("ofp_build_recipe", PKG / "build.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def create_for_build(binary: Path | None = None, compile_args: list[str] | None = None,
                     compiler: str | None = None) -> Path:
    """Sidecar for this package's CLI, as built by oneesan-gpu_amd/build.py."""
    if compile_args is None or compiler is None:
        recipe = _build_recipe()
        compile_args = recipe.HIP_FLAGS if compile_args is None else compile_args
        compiler = recipe.HIPCC if compiler is None else compiler
    binary = Path(binary) if binary else PKG / "lib" / "ofp_solver"
    csrc = PKG / "csrc"
    aux = [("library", PKG / "lib" / "libofp.so"),
           ("library-source-api", csrc / "ofp_api.hip"),
           ("library-source-cell", csrc / "ofp_cell.hip"),
           ("library-header-cell", csrc / "ofp_cell_impl.h"),
           ("library-header-internal", csrc / "ofp_internal.h"),
           ("abi-header", ROOT / "include" / "ofp.h")]
    return create(binary, source=csrc / "ofp_cli.cpp", compiler=Path(compiler),
                  compile_args=list(compile_args), auxiliary=aux)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="create or verify a solver build provenance sidecar")
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("create")
    c.add_argument("--binary", type=Path, default=None)
    v = sub.add_parser("verify")
    v.add_argument("provenance", type=Path)
    v.add_argument("--binary", type=Path, default=None)
    v.add_argument("--verify-sources", action="store_true")
    a = ap.parse_args(argv)
    try:
        if a.cmd == "create":
            print(create_for_build(a.binary))
        else:
            d = load(a.provenance, binary=a.binary, verify_sources=a.verify_sources)
            print(f"format={d['format']}\nbinary_sha256={d['binary_sha256']}\n"
                  f"dependencies={len(d['dependencies'])}\n"
                  f"auxiliary_dependencies={len(d.get('auxiliary_dependencies', []))}\nvalid=1")
    except (ValueError, OSError) as exc:
        print(f"INVALID: {exc}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
