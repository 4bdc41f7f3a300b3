"""Pin the CPU oracle (test infrastructure) before trusting it.

Checks, each chosen so a plausible mistake in the restatement fails it:
  * A007764 n=1..9 exact (scripts/test/golden_cpu.py:9-19) from BOTH oracles;
  * the reference's published residues that a CPU run reaches in seconds (n=9);
  * state counts equal the Motzkin closed form S(n) = M_{n+2} - M_n
    (Motzkin numbers from their own three-term recurrence, not from H_DP);
  * rank/unrank is a bijection onto legal words (brute-force enumeration);
  * forward (push) and pull (reverse-predecessor) cell steps agree on seeded
    random vectors at every cell (tests/factor_transfer_batch_test.cu pattern);
  * reverse predecessors = inverted forward relation with multiplicity
    (tests/gridfp_reverse_test.cpp:39-74) and the regression witness
    43348 -> 25600 at W=8 (tests/gridfp_reverse_test.cpp:24-37);
  * the oracle's include_horizontal / reverse enumeration equal the
    REFERENCE's own functions compiled from its headers (oracle/_ref, built by
    oracle/ref.mk when /root/reference is present), exhaustively for W=3..11.
"""
from __future__ import annotations

import ctypes
import itertools
from pathlib import Path

import numpy as np
import pytest

from conftest import ROOT, load_golden

REF_LIB = ROOT / "oracle" / "_ref" / "libref_gridfp.so"


def legal_words(width: int) -> list[int]:
    """All legal words (entered at height 1, never negative, end 0), brute force."""
    out = []
    for digits in itertools.product((0, 1, 2), repeat=width):
        h, ok = 1, True
        for d in digits:  # digits[0] is the highest position
            if d == 1:
                h -= 1
                if h < 0:
                    ok = False
                    break
            elif d == 2:
                h += 1
        if ok and h == 0:
            w = 0
            for d in digits:
                w = (w << 2) | d
            out.append(w)
    return out


def motzkin(m: int) -> int:
    a = [1, 1]
    for i in range(2, m + 1):
        a.append(((2 * i + 1) * a[i - 1] + (3 * i - 3) * a[i - 2]) // (i + 2))
    return a[m]


def test_a007764_exact_both_oracles(oracle):
    golden = load_golden("reference_residues.json")["exact_a007764"]
    for n in range(1, 10):
        want = int(golden[str(n)])
        assert oracle.exact_count(n) == want, n
        for p in (4294967291, 4294966997, 2147483647):
            assert oracle.solve(n, p, pull=False) == want % p, (n, p)
            assert oracle.solve(n, p, pull=True) == want % p, (n, p)


def test_exact_fixture_matches_grid_oracle(oracle):
    counts = load_golden("exact_counts.json")["counts"]
    for n in range(1, 12):
        c = int(counts[str(n)])
        for p in (4294967291, 1000003):
            assert oracle.solve(n, p) == c % p, (n, p)


def test_reference_n9_residues(oracle):
    for r in load_golden("reference_residues.json")["residues"]:
        if r["n"] <= 9:
            assert oracle.solve(r["n"], r["modulus"]) == r["residue"]


def test_state_counts_closed_form(oracle):
    lib = oracle.load()
    for n in range(1, 28):
        s = lib.or_main_count(n) + lib.or_blocked_count(n)
        assert s == motzkin(n + 2) - motzkin(n), n
    # n=27 counts quoted by the reference (docs/b300-hbm32.md:206-207)
    assert lib.or_main_count(27) == 385719506620
    assert lib.or_blocked_count(27) == 135015505407


@pytest.mark.parametrize("width", [1, 2, 3, 5, 8, 10])
def test_rank_unrank_bijection(oracle, width):
    lib = oracle.load()
    words = legal_words(width)
    assert len(words) == lib.or_dp(width, 1)
    for i, w in enumerate(sorted(words, key=lambda w: w)):
        pass
    # lexicographic order N<R<L from the high position = numeric order of the digit string
    ranks = [lib.or_rank(w, width) for w in words]
    assert ranks == list(range(len(words)))
    assert [lib.or_unrank(i, width) for i in range(len(words))] == words
    assert all(lib.or_is_legal(w, width) for w in words)


@pytest.mark.parametrize("n", [2, 4, 6, 8, 9])
def test_forward_equals_pull_every_cell(oracle, n):
    lib = oracle.load()
    nm, nb = lib.or_main_count(n), lib.or_blocked_count(n)
    rng = np.random.default_rng(20260905 + n)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    for mod in (2, 4294967291, 4294966997):
        for p in range(n, 0, -1):
            main = rng.integers(0, mod, nm, dtype=np.uint64).astype(np.uint32)
            blk = rng.integers(0, mod, nb, dtype=np.uint64).astype(np.uint32)
            outs = []
            for fn in (lib.or_cell_forward, lib.or_cell_pull):
                mo = np.zeros(nm, np.uint32)
                bo = np.zeros(nb, np.uint32)
                fn(n, p, mod, main.ctypes.data_as(u32p), blk.ctypes.data_as(u32p),
                   mo.ctypes.data_as(u32p), bo.ctypes.data_as(u32p))
                outs.append((mo, bo))
            assert np.array_equal(outs[0][0], outs[1][0]), (n, p, mod)
            assert np.array_equal(outs[0][1], outs[1][1]), (n, p, mod)


def _include(lib, fn, m, width, p):
    out, blocked = ctypes.c_uint64(), ctypes.c_int()
    ok = fn(m, width, p, ctypes.byref(out), ctypes.byref(blocked))
    return bool(ok), out.value, bool(blocked.value)


def test_reverse_inverts_forward_with_multiplicity(oracle):
    lib = oracle.load()
    buf = (ctypes.c_uint64 * 64)()
    for width in range(3, 10):
        main = legal_words(width)
        blocked = legal_words(width - 1)
        for p in range(2, width):
            expected: dict[int, list[int]] = {}
            for m in main:
                ok, t, is_blk = _include(lib, lib.or_include, m, width, p)
                if ok and is_blk:
                    expected.setdefault(t, []).append(m)
            for b in blocked:
                k = lib.or_reverse_block_preds(b, width, p, buf, 64)
                assert sorted(buf[i] for i in range(k)) == sorted(expected.get(b, [])), (width, p, b)
        boundary: dict[int, list[int]] = {}
        for m in main:
            ok, t, is_blk = _include(lib, lib.or_include, m, width, 1)
            assert not (ok and is_blk)
            if ok:
                boundary.setdefault(t, []).append(m)
        for t in main:
            k = lib.or_reverse_boundary_preds(t, width, buf, 64)
            assert sorted(buf[i] for i in range(k)) == sorted(boundary.get(t, [])), (width, t)


def test_regression_witness(oracle):
    # tests/gridfp_reverse_test.cpp:24-37 of the reference
    lib = oracle.load()
    w, is_blk = 43348, False
    for p in range(7, 0, -1):
        if is_blk:
            w, is_blk = lib.or_blocked_exclude(w, p), False
        elif p in (4, 2):
            ok, w, is_blk = _include(lib, lib.or_include, w, 8, p)
            assert ok
    ok, support, _ = _include(lib, lib.or_include, 1 << 14, 8, 6)
    assert not is_blk and w == support == 25600


@pytest.mark.skipif(not REF_LIB.exists(), reason="oracle/_ref not built (needs the reference tree)")
def test_oracle_transitions_equal_reference_headers(oracle):
    lib = oracle.load()
    ref = ctypes.CDLL(str(REF_LIB))
    u64, ip, u64p = ctypes.c_uint64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)
    ref.ref_include.argtypes = [u64, ctypes.c_int, ctypes.c_int, u64p, ip]
    ref.ref_include.restype = ctypes.c_int
    ref.ref_blocked_exclude.argtypes = [u64, ctypes.c_int]
    ref.ref_blocked_exclude.restype = u64
    for f in (ref.ref_reverse_block_preds,):
        f.argtypes = [u64, ctypes.c_int, ctypes.c_int, u64p, ctypes.c_int]
        f.restype = ctypes.c_int
    ref.ref_reverse_boundary_preds.argtypes = [u64, ctypes.c_int, u64p, ctypes.c_int]
    ref.ref_reverse_boundary_preds.restype = ctypes.c_int
    a, b = (ctypes.c_uint64 * 64)(), (ctypes.c_uint64 * 64)()
    for width in range(3, 12):
        main = legal_words(width)
        blocked = legal_words(width - 1)
        for p in range(1, width):
            for m in main:
                assert _include(lib, lib.or_include, m, width, p) == _include(ref, ref.ref_include, m, width, p), (width, p, m)
            for bw in blocked:
                assert lib.or_blocked_exclude(bw, p) == ref.ref_blocked_exclude(bw, p)
                if p >= 2:
                    ka = lib.or_reverse_block_preds(bw, width, p, a, 64)
                    kb = ref.ref_reverse_block_preds(bw, width, p, b, 64)
                    # same predecessors, same multiplicity (order may differ: sparse scan)
                    assert sorted(a[i] for i in range(ka)) == sorted(b[i] for i in range(kb)), (width, p, bw)
        for t in main:
            ka = lib.or_reverse_boundary_preds(t, width, a, 64)
            kb = ref.ref_reverse_boundary_preds(t, width, b, 64)
            assert sorted(a[i] for i in range(ka)) == sorted(b[i] for i in range(kb)), (width, t)


def test_ggcount_residues_fixture(oracle):
    """Residues written by tests/golden/make_ggcount_residues.py from the reference's
    third_party ggcount (an independent frontier DP) equal the Grid-FP oracle (n<=13)."""
    g = load_golden("ggcount_residues.json")["residues"]
    for r in g:
        if r["n"] <= 13:
            assert oracle.solve(r["n"], r["modulus"]) == r["residue"], r
