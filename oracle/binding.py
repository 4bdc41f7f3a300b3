"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_lib = None


def build() -> Path:
    subprocess.run(["make", "-C", str(HERE), "-s", "all"], check=True)
    return LIB


def load(auto_build: bool = True) -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not LIB.exists():
        if not auto_build:
            raise RuntimeError(f"{LIB} not built")
        build()
    lib = ctypes.CDLL(str(LIB))
    u32p = ctypes.POINTER(ctypes.c_uint32)
    u64 = ctypes.c_uint64
    ip = ctypes.POINTER(ctypes.c_int)
    sig = {
        "or_dp": (u64, [ctypes.c_int, ctypes.c_int]),
        "or_main_count": (u64, [ctypes.c_int]),
        "or_blocked_count": (u64, [ctypes.c_int]),
        "or_rank": (u64, [u64, ctypes.c_int]),
        "or_unrank": (u64, [u64, ctypes.c_int]),
        "or_is_legal": (ctypes.c_int, [u64, ctypes.c_int]),
        "or_include": (ctypes.c_int, [u64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u64), ip]),
        "or_blocked_exclude": (u64, [u64, ctypes.c_int]),
        "or_reverse_block_preds": (ctypes.c_int, [u64, ctypes.c_int, ctypes.c_int, ctypes.POINTER(u64), ctypes.c_int]),
        "or_reverse_boundary_preds": (ctypes.c_int, [u64, ctypes.c_int, ctypes.POINTER(u64), ctypes.c_int]),
        "or_cell_forward": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, u32p, u32p, u32p, u32p]),
        "or_cell_pull": (None, [ctypes.c_int, ctypes.c_int, ctypes.c_uint32, u32p, u32p, u32p, u32p]),
        "or_apply_cells": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, u32p, u32p, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_int]),
        "or_solve": (ctypes.c_uint32, [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, ip]),
        "or_num_threads": (ctypes.c_int, []),
        "or_remote_residue": (ctypes.c_uint32, [u64, u64, ctypes.c_int, ctypes.c_uint32]),
        "or_shard_cells": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint32, ctypes.c_int, u64, u64, u64, u64, u64,
                                          ctypes.c_int, ctypes.c_int, u32p, u32p]),
        "xdp_count_exact": (ctypes.c_int, [ctypes.c_int, ctypes.c_char_p, ctypes.c_int]),
        "xdp_count_mod": (ctypes.c_uint32, [ctypes.c_int, ctypes.c_uint32, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def solve(n: int, mod: int, pull: bool = True) -> int:
    lib = load()
    err = ctypes.c_int()
    r = lib.or_solve(n, mod, 1 if pull else 0, ctypes.byref(err))
    if err.value:
        raise ValueError(f"oracle solve failed n={n} mod={mod} err={err.value}")
    return int(r)


def exact_count(n: int) -> int:
    lib = load()
    buf = ctypes.create_string_buffer(64)
    if lib.xdp_count_exact(n, buf, 64):
        raise ValueError(f"exact DP supports 1 <= n <= 12, got {n}")
    return int(buf.value)


def count_mod_vertex_dp(n: int, mod: int) -> int:
    lib = load()
    err = ctypes.c_int()
    r = lib.xdp_count_mod(n, mod, ctypes.byref(err))
    if err.value:
        raise ValueError(f"vertex DP failed n={n} mod={mod}")
    return int(r)


def apply_cells(n: int, mod: int, main, blk, p_hi: int, p_lo: int, blocked_valid: bool,
                pull: bool = False) -> None:
    """main/blk: contiguous numpy uint32 1-D arrays for one modulus, updated in place."""
    lib = load()
    u32p = ctypes.POINTER(ctypes.c_uint32)
    rc = lib.or_apply_cells(n, mod, main.ctypes.data_as(u32p), blk.ctypes.data_as(u32p),
                            p_hi, p_lo, 1 if blocked_valid else 0, 1 if pull else 0)
    if rc:
        raise MemoryError("oracle apply_cells failed")


def shard_cells(n: int, mod: int, j: int, seed: int, b0: int, B: int, m0: int, M: int,
                p_hi: int, p_lo: int):
    """Windowed shard semantics (see or_shard_cells); returns (main_window, blocked_window)."""
    import numpy as np
    lib = load()
    mw = np.zeros(max(M, 1), np.uint32)
    bw = np.zeros(max(B, 1), np.uint32)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    rc = lib.or_shard_cells(n, mod, j, seed, b0, B, m0, M, p_hi, p_lo, mw.ctypes.data_as(u32p),
                            bw.ctypes.data_as(u32p))
    if rc:
        raise ValueError(f"or_shard_cells failed ({rc})")
    return mw[:M], bw[:B]
