"""ofp_solve_static, the entry point that allocates no device memory, against
ofp_solve through the C ABI: bit-identical residues on consecutive calls (the
second smaller, so stale contents of the static array would show), and a
refusal of states larger than the array."""
from __future__ import annotations

import ctypes

import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

MODS = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161, 4294967143, 4294967111]


def _static(lib, h, n, mods):
    arr = (ctypes.c_uint32 * len(mods))(*mods)
    out = (ctypes.c_uint32 * len(mods))()
    st = lib.ofp_solve_static(h, n, arr, len(mods), out)
    return st, list(out)


def test_static_solve_equals_solve():
    import ofp
    with ofp.Context() as ctx:
        lib, h = ctx._lib, ctx._h
        for n, mods in ((15, MODS), (11, MODS[:3]), (16, MODS[:2]), (9, MODS[:1])):
            print(f"case static n={n} k={len(mods)}", flush=True)
            st, got = _static(lib, h, n, mods)
            assert st == 0, lib.ofp_last_error(h)
            want, _ = ctx.solve(n, mods)
            assert got == want, (n, len(mods))
        exact = int(load_golden("exact_counts.json")["counts"]["12"])
        st, got = _static(lib, h, 12, MODS[:5])
        assert st == 0 and got == [exact % m for m in MODS[:5]]


def test_static_solve_refuses_large_states():
    import ofp
    with ofp.Context() as ctx:
        lib, h = ctx._lib, ctx._h
        for n, mods in ((18, MODS[:1]), (16, MODS[:3])):
            st, _ = _static(lib, h, n, mods)
            assert st == -5, (n, len(mods))  # OFP_E_MEMORY
            assert b"static array" in lib.ofp_last_error(h)
        st, _ = _static(lib, h, 9, [4294967291] * 17)
        assert st == -1
        # a refused call leaves the entry point usable
        st, got = _static(lib, h, 9, MODS[:2])
        assert st == 0 and got == ctx.solve(9, MODS[:2])[0]
