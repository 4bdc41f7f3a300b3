"""Checkpointed exact (CRT) runner.

    python -m ofp.crt_runner N [--binary PATH] [--work-dir DIR] [--max-runs K]

Behaviour follows the reference's scripts/solve/solve_b300_exact_batch.py:70-264:
bound -> prime prefix -> strict checkpoint load (bound to the solver binary's
SHA-256, one runner per work dir via flock) -> ONE child process with every
missing modulus, whose result lines are parsed as they stream, each residue
checkpointed durably before the next is read -> CRT -> exact.txt +
exact_manifest.json.  The child is the drop-in CLI (oneesan-gpu_amd/lib/
ofp_solver), which computes all the moduli in one interleaved sweep.

Fixes relative to the reference: the reference's batch runner calls an
undefined solver_fingerprint() (solve_b300_exact_batch.py:106) and raises
NameError at start; here the checkpoint identity (solver_identity) is used.

Build provenance: when the solver binary has a sidecar next to it
(<binary>.provenance.json, written by oneesan-gpu_amd/build.py), it is checked
against the binary at start and copied into the work directory as
build_provenance.json, which the manifest names and hashes.

Exit codes: 0 complete (or partial with --max-runs), 2 partial, others are
errors (SystemExit with a message).
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

from . import provenance
from .exact_safety import (
    acquire_workdir_lock, crt_pair, durable_atomic_write, load_checkpoint, save_checkpoint,
    solver_identity, validate_exact_reconstruction, validate_residue, write_exact_result,
)
from .path_bound import primes_for_bound, simple_path_upper_bound
from .solver_output import parse_result_line

PKG = Path(__file__).resolve().parents[1]
DEFAULT_BINARY = PKG / "lib" / "ofp_solver"


def reconstruct(prefix: list[int], residues: dict[int, dict]) -> tuple[int, int, float, int]:
    """CRT over the longest contiguous prefix of known residues."""
    x, m, wall, used = 0, 1, 0.0, 0
    for p in prefix:
        if p not in residues:
            break
        r = int(residues[p]["residue"])
        validate_residue(p, r, source=f"checkpoint modulus {p}")
        x, m = crt_pair(x, m, r, p)
        wall += float(residues[p].get("wall_s", 0.0))
        used += 1
    return x, m, wall, used


def finish(work: Path, n: int, path_bound: int, prefix: list[int], residues: dict[int, dict],
           strip_partition: list[int], identity: dict[str, str], checkpoint: Path,
           out=sys.stdout, build_provenance: Path | None = None) -> int:
    x, m, wall, used = reconstruct(prefix, residues)
    if m <= path_bound:
        raise RuntimeError("finish called before CRT capacity reached")
    # smallest prefix whose product exceeds the bound
    k, mm = 0, 1
    for p in prefix[:used]:
        k += 1
        mm *= p
        if mm > path_bound:
            break
    used_moduli = prefix[:k]
    x, m, wall, _ = reconstruct(used_moduli, residues)
    validate_exact_reconstruction(x, m, path_bound, [(p, residues[p]["residue"]) for p in used_moduli])
    exact_path, manifest = write_exact_result(
        work, n=n, exact=x, path_bound=path_bound, modulus_product=m, used_moduli=used_moduli,
        residues=residues, identity=identity, checkpoint_path=checkpoint, total_wall_s=wall,
        strip_partition=strip_partition, build_provenance=build_provenance)
    print(f"n={n}\nexact={x}\nbound_bits={path_bound.bit_length()}\nmodulus_bits={m.bit_length()}\n"
          f"primes_used={k}\nsolver_wall_s_sum={wall:.9f}\nresult_file={exact_path}\n"
          f"manifest_file={manifest}", file=out)
    return 0


def run(n: int, binary: Path, work: Path, *, target_mib: int = 16384, max_window: int = 14,
        gpus: int = 1, max_runs: int = 0, env: dict | None = None, out=sys.stdout, err=sys.stderr) -> int:
    path_bound, strips = simple_path_upper_bound(n)
    prefix = primes_for_bound(path_bound)
    bits = path_bound.bit_length()
    print(f"rigorous path bound: <={path_bound} ({bits} bits), strips={strips}, CRT primes={len(prefix)}",
          file=err)
    binary = Path(binary)
    if not binary.exists():
        raise SystemExit(f"solver binary not found: {binary}")
    binary = binary.resolve()
    work = Path(work)
    work.mkdir(parents=True, exist_ok=True)
    try:
        lock = acquire_workdir_lock(work)
    except RuntimeError as exc:
        raise SystemExit(str(exc)) from exc
    try:
        checkpoint = work / "checkpoint.json"
        identity = solver_identity(binary)
        print(f"solver identity: sha256={identity['binary_sha256']}", file=err)
        prov = None
        sidecar = Path(str(binary) + provenance.SIDECAR_SUFFIX)
        if sidecar.is_file():
            try:
                provenance.load(sidecar, binary=binary)
            except ValueError as exc:
                raise SystemExit(f"solver build provenance does not match the binary: {exc}") from exc
            prov = work / "build_provenance.json"
            durable_atomic_write(prov, sidecar.read_text(encoding="utf-8"))
        try:
            residues = load_checkpoint(checkpoint, n=n, identity=identity)
        except ValueError as exc:
            raise SystemExit(str(exc)) from exc
        unexpected = sorted(set(residues) - set(prefix))
        if unexpected:
            raise SystemExit(f"checkpoint contains unexpected moduli: {unexpected}")
        _, m, _, _ = reconstruct(prefix, residues)
        if m > path_bound:
            return finish(work, n, path_bound, prefix, residues, strips, identity, checkpoint, out, prov)
        missing = [p for p in prefix if p not in residues]
        if max_runs:
            missing = missing[:max_runs]
        cmd = [str(binary), str(n), str(target_mib), str(max_window), str(gpus), *map(str, missing)]
        print(f"batch run: {len(missing)} residue(s), first={missing[0]}, last={missing[-1]}", file=err, flush=True)
        proc = subprocess.Popen(cmd, text=True, stdout=subprocess.PIPE, stderr=None,
                                env=env if env is not None else os.environ.copy(), bufsize=1)
        assert proc.stdout is not None
        seen: set[int] = set()

        def abort(msg: str):
            proc.terminate()
            proc.wait()
            raise SystemExit(msg)
        for line in proc.stdout:
            out.write(line)
            out.flush()
            try:
                res = parse_result_line(line)
            except (ValueError, OverflowError) as exc:
                abort(f"invalid batch solver result line: {line.strip()}: {exc}")
            if res is None:
                continue
            if res.n != n:
                abort(f"batch solver returned n={res.n}, expected n={n}")
            if res.modulus not in missing:
                abort(f"batch solver returned unexpected modulus {res.modulus}")
            if res.modulus in seen:
                abort(f"batch solver returned duplicate modulus {res.modulus}")
            try:
                validate_residue(res.modulus, res.residue, source=f"batch solver modulus {res.modulus}")
            except ValueError as exc:
                abort(str(exc))
            residues[res.modulus] = {"residue": res.residue, "wall_s": res.wall_s}
            seen.add(res.modulus)
            save_checkpoint(checkpoint, n=n, identity=identity, residues=residues)
            _, mm, _, _ = reconstruct(prefix, residues)
            print(f"checkpoint p={res.modulus}: contiguous CRT bits={mm.bit_length()} / bound_bits={bits}",
                  file=err, flush=True)
        rc = proc.wait()
        if rc != 0:
            raise SystemExit(f"batch solver failed, rc={rc}; completed this invocation={len(seen)}/{len(missing)}")
        if seen != set(missing):
            raise SystemExit(f"batch solver exited without residues for: {sorted(set(missing) - seen, reverse=True)}")
        _, m, _, _ = reconstruct(prefix, residues)
        if m > path_bound:
            return finish(work, n, path_bound, prefix, residues, strips, identity, checkpoint, out, prov)
        print(f"partial: CRT bits={m.bit_length()} bound_bits={bits}; cached={len(residues)}", file=err)
        return 0 if max_runs else 2
    finally:
        lock.close()


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="checkpointed multi-modulus CRT runner")
    ap.add_argument("n", nargs="?", type=int, default=22)
    ap.add_argument("--binary", default=str(DEFAULT_BINARY))
    ap.add_argument("--target-mib", type=int, default=16384)
    ap.add_argument("--max-window", type=int, default=14)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--work-dir", default=None)
    ap.add_argument("--max-runs", type=int, default=0)
    a = ap.parse_args(argv)
    if not 1 <= a.n <= 27:
        ap.error("n must be in 1..27")
    if a.target_mib < 1 or a.max_window < 1 or not 0 <= a.gpus <= 8 or a.max_runs < 0:
        ap.error("invalid --target-mib/--max-window/--gpus/--max-runs")
    work = Path(a.work_dir) if a.work_dir else Path.cwd() / "work" / f"exact_n{a.n}"
    return run(a.n, Path(a.binary), work, target_mib=a.target_mib, max_window=a.max_window,
               gpus=a.gpus, max_runs=a.max_runs)


if __name__ == "__main__":
    raise SystemExit(main())
