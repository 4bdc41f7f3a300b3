"""GPU parity of the HIP cell sweep against the CPU oracle (through the C ABI).

Pattern follows the reference's own GPU tests (tests/factor_transfer_batch_test.cu:44-96):
seeded random residues in [0, p) on every main and blocked word, one window of
cells, every coefficient compared with the CPU forward recurrence, moduli
{2, 4294967291, 4294966997}.  End-to-end residues are compared with the
reference's published values (tests/golden/reference_residues.json) and with
the independent vertex/component DP.  The bar is bit-exact.
"""
from __future__ import annotations

import numpy as np
import pytest

from conftest import load_golden

pytestmark = pytest.mark.gpu

MODS3 = [2, 4294967291, 4294966997]


def _oracle_cells(oracle, n, mod, main, blk, p_hi, p_lo, blocked_valid):
    m = np.ascontiguousarray(main.copy())
    b = np.ascontiguousarray(blk.copy())
    oracle.apply_cells(n, mod, m, b, p_hi, p_lo, blocked_valid)
    return m, b


@pytest.mark.parametrize("n", [2, 3, 5, 7, 9])
def test_every_cell_random_vectors(gpu_ctx, oracle, n):
    """Each single cell p on random full vectors, k=3 moduli interleaved."""
    import ofp
    nm, nb = ofp.state_counts(n)
    W = n + 1
    rng = np.random.default_rng(0x62C31F05 + n)
    for p in range(W - 1, 0, -1):
        for blocked_valid in (True, False):
            main = np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in MODS3], 1).astype(np.uint32)
            blk = np.stack([rng.integers(0, m, nb, dtype=np.uint64) for m in MODS3], 1).astype(np.uint32)
            g_main, g_blk = main.copy(), blk.copy()
            print(f"case n={n} p={p} blocked_valid={blocked_valid}", flush=True)
            gpu_ctx.apply_cells(n, MODS3, g_main, g_blk, p, p, blocked_valid)
            for j, mod in enumerate(MODS3):
                em, eb = _oracle_cells(oracle, n, mod, main[:, j], blk[:, j], p, p, blocked_valid)
                assert np.array_equal(g_main[:, j], em), (n, p, mod, "main")
                assert np.array_equal(g_blk[:, j], eb), (n, p, mod, "blocked")


@pytest.fixture(scope="module", params=[(0, 0), (1, 0), (1, 1), (1, 2), (1, 3), (1, 4)],
                ids=["tiles64", "statemajor-early", "statemajor-late", "statemajor-uniform",
                     "statemajor-early-retire", "statemajor-sorted"])
def layout_ctx(request):
    import ofp
    layout, variant = request.param
    ctx = ofp.Context(layout=layout)
    ctx.set_variant(variant)
    ctx.set_pair_passes(0)  # these cases test the one-launch-per-cell kernels
    yield ctx
    ctx.close()


@pytest.mark.parametrize("n,k", [(8, 1), (9, 14), (10, 17), (11, 5), (7, 3)])
def test_whole_row_windows(layout_ctx, oracle, n, k):
    """A window of consecutive cells (two windows covering a row), k moduli, both layouts."""
    import ofp
    gpu_ctx = layout_ctx
    nm, nb = ofp.state_counts(n)
    W = n + 1
    primes = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161,
              4294967143, 4294967111, 4294967087, 4294967029, 4294966997, 4294966981,
              4294966943, 4294966927, 4294966909, 4294966877, 4294966829]
    mods = (primes + [2, 3, 65537])[:k] if k <= 17 else None
    rng = np.random.default_rng(1234 + n * 31 + k)
    half = (W - 1) // 2 + 1
    for p_hi, p_lo in ((W - 1, half), (half - 1, 1)):
        main = np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
        blk = np.stack([rng.integers(0, m, nb, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
        g_main, g_blk = main.copy(), blk.copy()
        print(f"case n={n} k={k} window={p_hi}..{p_lo}", flush=True)
        gpu_ctx.apply_cells(n, mods, g_main, g_blk, p_hi, p_lo, True)
        for j in sorted({0, k // 2, k - 1}):
            em, eb = _oracle_cells(oracle, n, mods[j], main[:, j], blk[:, j], p_hi, p_lo, True)
            assert np.array_equal(g_main[:, j], em), (n, k, j, p_hi, p_lo)
            assert np.array_equal(g_blk[:, j], eb), (n, k, j, p_hi, p_lo)


def test_small_n_against_exact_counts(gpu_ctx, oracle):
    """n=1..12: GPU residues equal the exact vertex/component DP count mod p."""
    mods = [4294967291, 4294966997, 2147483647, 2, 3, 1000003]
    for n in range(1, 13):
        print(f"case n={n}", flush=True)
        got, _ = gpu_ctx.solve(n, mods)
        exact = int(load_golden("exact_counts.json")["counts"][str(n)])
        assert got == [exact % m for m in mods], n


def test_reference_residues(layout_ctx):
    """Every published reference residue with n <= 21 (n=22/23 are bench-sized), both layouts."""
    gpu_ctx = layout_ctx
    g = load_golden("reference_residues.json")
    by_n: dict[int, list[tuple[int, int]]] = {}
    for r in g["residues"]:
        if r["n"] <= 21:
            by_n.setdefault(r["n"], []).append((r["modulus"], r["residue"]))
    for n, pairs in sorted(by_n.items()):
        print(f"case reference n={n}", flush=True)
        got, _ = gpu_ctx.solve(n, [m for m, _ in pairs])
        assert got == [r for _, r in pairs], n


def test_batched_equals_one_at_a_time(gpu_ctx):
    """k moduli interleaved give the same residues as k separate one-modulus runs."""
    mods = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161,
            4294967143, 4294967111, 4294967087, 4294967029, 4294966997, 4294966981,
            4294966943, 4294966927, 4294966909, 4294966877, 4294966829, 4294966813]
    n = 14
    batched, _ = gpu_ctx.solve(n, mods)
    single = [gpu_ctx.solve(n, [m])[0][0] for m in mods]
    assert batched == single


def test_n12_config_residue(gpu_ctx, oracle):
    """Config 1 (n=12, p=4294967291): GPU == Grid-FP oracle == vertex DP."""
    got, _ = gpu_ctx.solve(12, [4294967291])
    assert got[0] == oracle.solve(12, 4294967291)
    assert got[0] == int(load_golden("exact_counts.json")["counts"]["12"]) % 4294967291


def test_ggcount_residues(gpu_ctx):
    """n=10..19 against the reference's third_party ggcount (independent frontier DP)."""
    rows = load_golden("ggcount_residues.json")["residues"]
    by_n: dict[int, list[tuple[int, int]]] = {}
    for r in rows:
        by_n.setdefault(r["n"], []).append((r["modulus"], r["residue"]))
    for n, pairs in sorted(by_n.items()):
        print(f"case ggcount n={n}", flush=True)
        got, _ = gpu_ctx.solve(n, [m for m, _ in pairs])
        assert got == [r for _, r in pairs], n


def test_height_caps_match_whole_state_sweep(gpu_ctx):
    """Capped first and last rows give the residues of the uncapped sweep, also when
    the rows are swept in several calls (the row count carries across calls)."""
    mods = [4294967291, 4294966997, 2147483647]
    try:
        for n in (3, 5, 8, 11, 14, 17):
            W = n + 1
            print(f"case caps n={n}", flush=True)
            gpu_ctx.set_height_caps(False)
            whole, _ = gpu_ctx.solve(n, mods)
            gpu_ctx.set_height_caps(100)
            every, _ = gpu_ctx.solve(n, mods)
            assert every == whole, n
            gpu_ctx.set_height_caps(True)
            capped, st = gpu_ctx.solve(n, mods)
            assert capped == whole, n
            if n >= 8:
                assert st["capped_launches"] > 0 and st["full_launches"] > 0
            gpu_ctx.prepare(n, mods)
            gpu_ctx.row_sweep(2)
            gpu_ctx.row_sweep(W - 3)
            gpu_ctx.row_sweep(1)
            assert gpu_ctx.read_main(0, len(mods)) == whole, n
        # one modulus: the default threshold is lower (75 %)
        for n in (11, 17):
            print(f"case caps n={n} k=1", flush=True)
            gpu_ctx.set_height_caps(False)
            whole, _ = gpu_ctx.solve(n, mods[:1])
            gpu_ctx.set_height_caps(True)
            capped, st = gpu_ctx.solve(n, mods[:1])
            assert capped == whole and st["capped_launches"] > 0, n
    finally:
        gpu_ctx.set_height_caps(True)


@pytest.mark.parametrize("n", [16, 23])
def test_sorted_cells_same_residues(gpu_ctx, n):
    """The sorted assignment on the upper cells changes no residue: n=16 with several
    moduli (32-bit ranks) and n=23 with one (64-bit ranks, the reference's residue)."""
    mods = [4294967291, 4294966997, 2147483647, 65537] if n == 16 else [4294967291]
    try:
        gpu_ctx.set_pair_passes(0)  # the sorted kernel runs on whole-state cells launched one by one
        gpu_ctx.set_sorted_cells(0)
        if n == 23:
            want = [2762394459]
        else:
            want, _ = gpu_ctx.solve(n, mods)
        for sf in ((1, 6, 12) if n == 16 else (6,)):
            print(f"case sorted cells from {sf}, n={n}", flush=True)
            gpu_ctx.set_sorted_cells(sf)
            got, _ = gpu_ctx.solve(n, mods)
            assert got == want, (n, sf)
    finally:
        gpu_ctx.set_sorted_cells(-1)  # the library's default
        gpu_ctx.set_pair_passes(1)


def test_xcd_pieces_same_residues(gpu_ctx):
    """Dealing the chunks to the XCDs in pieces of any size changes no residue."""
    mods = [4294967291, 4294966997, 2147483647, 65537]
    try:
        gpu_ctx.set_xcd_pieces(-1)
        want, _ = gpu_ctx.solve(15, mods)
        for shift in (0, 1, 4, 9):
            print(f"case xcd pieces 2^{shift}", flush=True)
            gpu_ctx.set_xcd_pieces(shift)
            got, _ = gpu_ctx.solve(15, mods)
            assert got == want, shift
    finally:
        gpu_ctx.set_xcd_pieces(-1)


@pytest.mark.parametrize("n", [22, 23])
def test_largest_reference_residues(gpu_ctx, n):
    """n=22 (the bench size; 32-bit index path) and n=23 (64-bit index path, 7.8e9 states)
    one-modulus solves against the residues the reference publishes."""
    g = load_golden("reference_residues.json")
    pairs = [(r["modulus"], r["residue"]) for r in g["residues"] if r["n"] == n]
    assert pairs
    print(f"case reference n={n} (whole solve, one modulus)", flush=True)
    got, st = gpu_ctx.solve(n, [pairs[0][0]])
    assert got == [pairs[0][1]], (n, got)


PRIMES16 = [4294967291, 4294967279, 4294967231, 4294967197, 4294967189, 4294967161, 4294967143,
            4294967111, 4294967087, 4294967029, 4294966997, 4294966981, 4294966943, 4294966927,
            4294966909, 4294966877]


@pytest.mark.parametrize("k", [1, 8, 14])
@pytest.mark.parametrize("sorted_cells", [False, True], ids=["unsorted", "sorted"])
def test_64bit_ranks_every_cell(gpu_ctx, oracle, k, sorted_cells):
    """The 64-bit-rank kernels (the path n >= 23 takes) forced at small n: every
    cell on random vectors, k moduli interleaved, unsorted (cell_kernel<k, uint64>)
    and sorted (cell_sorted_kernel<k, uint64>, the default of n >= 23 with 8 or more
    moduli), against the CPU forward recurrence."""
    import ofp
    mods = ([2] + PRIMES16)[:k] if k > 1 else [4294967291]
    try:
        gpu_ctx.set_index_width(64)
        gpu_ctx.set_sorted_cells(1 if sorted_cells else 0)
        for n in (7, 10, 11):
            nm, nb = ofp.state_counts(n)
            W = n + 1
            rng = np.random.default_rng(0x64B17 + 97 * n + k)
            cells = range(W - 1, 0, -1) if n < 11 else (W - 1, W // 2, 3, 2, 1)
            for p in cells:
                main = np.stack([rng.integers(0, m, nm, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                blk = np.stack([rng.integers(0, m, nb, dtype=np.uint64) for m in mods], 1).astype(np.uint32)
                g_main, g_blk = main.copy(), blk.copy()
                print(f"case 64-bit n={n} k={k} p={p} sorted={sorted_cells}", flush=True)
                gpu_ctx.apply_cells(n, mods, g_main, g_blk, p, p, True)
                for j in sorted({0, k // 2, k - 1}):
                    em, eb = _oracle_cells(oracle, n, mods[j], main[:, j], blk[:, j], p, p, True)
                    assert np.array_equal(g_main[:, j], em), (n, k, p, j, "main")
                    assert np.array_equal(g_blk[:, j], eb), (n, k, p, j, "blocked")
    finally:
        gpu_ctx.set_index_width(0)
        gpu_ctx.set_sorted_cells(-1)


def test_64bit_ranks_whole_solves(gpu_ctx):
    """Whole solves with 64-bit ranks forced (capped, unsorted and sorted kernels):
    residues equal the 32-bit path's and the exact counts."""
    mods = PRIMES16[:9]
    try:
        for n in (9, 12, 15):
            gpu_ctx.set_index_width(0)
            want, _ = gpu_ctx.solve(n, mods)
            gpu_ctx.set_index_width(64)
            print(f"case 64-bit whole solve n={n} k={len(mods)}", flush=True)
            got, st = gpu_ctx.solve(n, mods)
            assert got == want, n
        exact = int(load_golden("exact_counts.json")["counts"]["12"])
        gpu_ctx.set_index_width(64)
        got, _ = gpu_ctx.solve(12, mods)
        assert got == [exact % m for m in mods]
    finally:
        gpu_ctx.set_index_width(0)


def test_n23_two_moduli():
    """n=23 (64-bit ranks, 7.8e9 states) with two moduli interleaved: the first
    residue is the reference's (docs/b300-hbm32.md:382-412), the second equals a
    one-modulus solve."""
    import ofp
    mods = [4294967291, 4294967279]
    with ofp.Context() as ctx:
        print("case n=23 k=2 interleaved", flush=True)
        got, _ = ctx.solve(23, mods)
        assert got[0] == 2762394459
        print("case n=23 k=1 second modulus", flush=True)
        one, _ = ctx.solve(23, mods[1:])
        assert one == got[1:]
