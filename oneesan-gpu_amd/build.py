"""Build recipe for the MI355X frontier solver (gfx950).

Produces, in-tree:
  oneesan-gpu_amd/lib/libofp.so    the C-ABI library (HIP kernels + host runtime)
  oneesan-gpu_amd/lib/ofp_solver   the CLI (reference argv / result-line contract)
  oneesan-gpu_amd/lib/ofp_solver.provenance.json  its build provenance sidecar

hipcc cross-compiles for gfx950 without a GPU.  Every HIP translation unit is
built with packed-FP32 code generation disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"

ARCH = os.environ.get("OFP_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

LIB_SOURCES = ["ofp_cell.hip", "ofp_pair.hip", "ofp_exchange.hip", "ofp_api.hip"]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(target: Path, deps: list[Path]) -> bool:
    if not target.exists():
        return True
    t = target.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def build(force: bool = True, jobs: int = 4) -> Path:
    """Compile libofp.so and the CLI.  Always recompiles when force is set."""
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    LIB.mkdir(exist_ok=True)
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    headers = list(CSRC.glob("*.h")) + list(INCLUDE.glob("*.h"))
    objs: list[Path] = []
    procs = []
    for src in LIB_SOURCES:
        s = CSRC / src
        o = objdir / (s.stem + ".o")
        objs.append(o)
        if force or _stale(o, [s] + headers):
            cmd = [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)]
            print("+", " ".join(cmd), flush=True)
            procs.append(subprocess.Popen(cmd))
    for p in procs:
        if p.wait() != 0:
            raise subprocess.CalledProcessError(p.returncode, p.args)
    so = LIB / "libofp.so"
    tmp = LIB / "libofp.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, so)
    cli_src = CSRC / "ofp_cli.cpp"
    if cli_src.exists():
        cli = LIB / "ofp_solver"
        _run([HIPCC, *HIP_FLAGS, "-x", "c++", str(cli_src), f"-L{LIB}", "-lofp",
              f"-Wl,-rpath,$ORIGIN", "-o", str(cli)])
        # provenance sidecar lib/ofp_solver.provenance.json (what the binary was built from)
        sys.path.insert(0, str(PKG))
        from ofp import provenance
        provenance.create_for_build(cli, compile_args=HIP_FLAGS, compiler=HIPCC)
    return so


if __name__ == "__main__":
    build(force="--no-force" not in sys.argv)
