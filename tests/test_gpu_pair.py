"""The pair pass on the GPU (two cells per out-of-place launch), through the C
ABI, bit-exact: windows of cells on seeded random vectors against the CPU
oracle's cells, and whole solves against the one-launch-per-cell sweep and the
reference's published residues."""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

PRIMES = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161, 4294967143,
          4294967111, 4294967087, 4294967029, 4294966997, 4294966981, 4294966943, 4294966927]


@pytest.fixture(scope="module")
def pair_ctx():
    import ofp
    ctx = ofp.Context()
    ctx.set_pair_passes(True, 0)
    yield ctx
    ctx.close()


def _oracle_cells(oracle, n, mod, main, blk, p_hi, p_lo, blocked_valid):
    m = np.ascontiguousarray(main.copy())
    b = np.ascontiguousarray(blk.copy())
    oracle.apply_cells(n, mod, m, b, p_hi, p_lo, blocked_valid)
    return m, b


@pytest.mark.parametrize("k", [1, 3, 8, 14])
def test_pair_windows_random_vectors(pair_ctx, oracle, k):
    """Every pair of cells (p, p-1) and odd windows (a pair plus one cell) on random
    vectors, blocked input on and off, k moduli (14 = launches of 8 and 6)."""
    import ofp
    mods = ([2] + PRIMES)[:k] if k > 1 else [4294967291]
    for n in (2, 3, 5, 8, 11):
        nm, nb = ofp.state_counts(n)
        W = n + 1
        rng = np.random.default_rng(0x9A12 + 31 * n + k)
        windows = [(p, p - 1) for p in range(W - 1, 1, -1)] + ([(W - 1, 1)] if n >= 3 else [])
        if n >= 5:
            windows.append((W - 2, W - 4))  # three cells: a pair, then one cell in place
        for p_hi, p_lo in windows:
            for bv in (True, False):
                main = np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                blk = np.stack([rng.integers(0, m, nb, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                g_main, g_blk = main.copy(), blk.copy()
                print(f"case pair n={n} k={k} window={p_hi}..{p_lo} blocked={bv}", flush=True)
                pair_ctx.apply_cells(n, mods, g_main, g_blk, p_hi, p_lo, bv)
                for j in sorted({0, k // 2, k - 1}):
                    em, eb = _oracle_cells(oracle, n, mods[j], main[:, j], blk[:, j], p_hi, p_lo, bv)
                    assert np.array_equal(g_main[:, j], em), (n, k, p_hi, p_lo, bv, j, "main")
                    assert np.array_equal(g_blk[:, j], eb), (n, k, p_hi, p_lo, bv, j, "blocked")


def test_pair_solves_equal_cell_solves(pair_ctx, gpu_ctx):
    """Whole solves with pair passes (height-capped rows still per cell) equal the
    per-cell sweep's, for even and odd cell counts and several moduli counts."""
    for n, k in ((3, 2), (4, 5), (9, 1), (12, 8), (15, 14), (16, 3), (17, 8)):
        mods = PRIMES[:k]
        print(f"case pair solve n={n} k={k}", flush=True)
        want, _ = gpu_ctx.solve(n, mods)
        got, st = pair_ctx.solve(n, mods)
        assert got == want, (n, k)
    exact = int(load_golden("exact_counts.json")["counts"]["12"])
    got, _ = pair_ctx.solve(12, [4294967291, 4294966997, 2147483647])
    assert got == [exact % m for m in (4294967291, 4294966997, 2147483647)]


def test_pair_masks_same_residues(pair_ctx, gpu_ctx):
    """Any placement of pair passes (cells left out of pairs run one launch each)
    gives the same residues."""
    mods = PRIMES[:5]
    try:
        for n in (12, 15):
            gpu_ctx.set_pair_passes(0)
            want, _ = gpu_ctx.solve(n, mods)
            for mask in (0, 0xFFFFFFFF, 0x55555555, 0xAAAAAAAA, 0x1234, 0x18):
                print(f"case pair mask {mask:#x} n={n}", flush=True)
                pair_ctx.set_pair_passes(1, mask)
                got, _ = pair_ctx.solve(n, mods)
                assert got == want, (n, hex(mask))
            pair_ctx.set_pair_passes(8, 0)  # at most 8 moduli per launch
            got, _ = pair_ctx.solve(n, PRIMES[:11])
            gpu_ctx.set_pair_passes(0)
            assert got == gpu_ctx.solve(n, PRIMES[:11])[0], n
    finally:
        pair_ctx.set_pair_passes(1, 0)
        gpu_ctx.set_pair_passes(1, 0)


def test_pair_reference_residues(pair_ctx):
    """The reference's published residues up to n = 21 with pair passes."""
    g = load_golden("reference_residues.json")
    by_n: dict[int, list[tuple[int, int]]] = {}
    for r in g["residues"]:
        if r["n"] <= 21:
            by_n.setdefault(r["n"], []).append((r["modulus"], r["residue"]))
    for n, pairs in sorted(by_n.items()):
        print(f"case pair reference n={n}", flush=True)
        got, _ = pair_ctx.solve(n, [m for m, _ in pairs])
        assert got == [r for _, r in pairs], n


def test_pair_rows_split_over_calls(pair_ctx):
    """Rows swept over several calls (the copies trade places each pass) give the
    whole solve's residues."""
    mods = PRIMES[:6]
    n = 14
    want, _ = pair_ctx.solve(n, mods)
    pair_ctx.prepare(n, mods)
    pair_ctx.row_sweep(3)
    pair_ctx.row_sweep(7)
    pair_ctx.row_sweep(n + 1 - 10)
    assert pair_ctx.read_main(0, len(mods)) == want


@pytest.mark.parametrize("k", [1, 8, 14])
def test_pair_windows_64bit_ranks(pair_ctx, oracle, k):
    """The pair pass with 64-bit ranks (the path n >= 23 takes), forced at small n:
    every pair of cells on random vectors, blocked input on and off, against the
    oracle's cells."""
    import ofp
    mods = ([2] + PRIMES)[:k] if k > 1 else [4294967291]
    try:
        pair_ctx.set_index_width(64)
        for n in (3, 7, 11):
            nm, nb = ofp.state_counts(n)
            W = n + 1
            rng = np.random.default_rng(0x64A1 + 31 * n + k)
            for p_hi in range(W - 1, 1, -1):
                for bv in (True, False):
                    main = np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                    blk = np.stack([rng.integers(0, m, nb, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                    g_main, g_blk = main.copy(), blk.copy()
                    print(f"case 64-bit pair n={n} k={k} window={p_hi}..{p_hi - 1} blocked={bv}", flush=True)
                    pair_ctx.apply_cells(n, mods, g_main, g_blk, p_hi, p_hi - 1, bv)
                    for j in sorted({0, k // 2, k - 1}):
                        em, eb = _oracle_cells(oracle, n, mods[j], main[:, j], blk[:, j], p_hi, p_hi - 1, bv)
                        assert np.array_equal(g_main[:, j], em), (n, k, p_hi, bv, j, "main")
                        assert np.array_equal(g_blk[:, j], eb), (n, k, p_hi, bv, j, "blocked")
    finally:
        pair_ctx.set_index_width(0)
