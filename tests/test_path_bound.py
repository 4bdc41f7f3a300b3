"""CRT prime table and checkerboard-strip bound (oneesan-gpu_amd/ofp/path_bound.py).

Known-answer checks follow the reference's tests/test_path_bound.py:88-165 and
tests/test_crt_primes.py, plus the check the reference lacks: the table is the
48 CONSECUTIVE largest primes below 2^32.
"""
from __future__ import annotations

import itertools

from ofp import path_bound as pb

N27_BOUND = int(
    "31732427633797389964407887052573851640105323179333844527763421102211579188310190597"
    "41290075655087412312934258526184013896401019031236450810716892700623227706778327992"
    "7977876779029754107293528")


def brute_no_checkerboard(h: int, w: int) -> int:
    total = 0
    for mask in range(1 << (h * w)):
        ok = True
        for r in range(h - 1):
            for c in range(w - 1):
                a = (mask >> (r * w + c)) & 1
                b = (mask >> ((r + 1) * w + c)) & 1
                x = (mask >> (r * w + c + 1)) & 1
                y = (mask >> ((r + 1) * w + c + 1)) & 1
                if a != b and x != y and a != x:
                    ok = False
        total += ok
    return total


def test_strip_count_equals_brute_force():
    for h, w in [(1, 1), (1, 4), (2, 2), (2, 4), (3, 3), (4, 3), (3, 4)]:
        assert pb.strip_count(h, w) == brute_no_checkerboard(h, w), (h, w)


def test_partition_is_minimal_for_small_n():
    for n in range(1, 8):
        counts = {h: pb.strip_count(h, n) for h in range(1, min(4, n) + 1)}

        def products(m):
            if m == 0:
                yield 1
                return
            for h, c in counts.items():
                if h <= m:
                    for rest in products(m - h):
                        yield c * rest
        bound, parts = pb.simple_path_upper_bound(n, max_strip_height=4)
        assert bound == min(products(n))
        assert sum(parts) == n


def test_bound_dominates_true_counts():
    a007764 = [2, 12, 184, 8512, 1262816, 575780564, 789360053252, 3266598486981642,
               41044208702632496804]
    for n, a in enumerate(a007764, start=1):
        assert pb.simple_path_upper_bound(n)[0] >= a


def test_n27_known_answer():
    bound, parts = pb.simple_path_upper_bound(27)
    assert parts == [9, 9, 9]
    assert bound.bit_length() == 633
    assert bound == N27_BOUND
    assert len(pb.primes_for_bound(bound)) == 20


def test_prime_counts_for_configs():
    assert len(pb.primes_for_n(4)) == 1
    assert len(pb.primes_for_n(22)) == 14
    assert pb.simple_path_upper_bound(22)[0].bit_length() == 423


def test_primes_are_consecutive_largest_below_2_32():
    lo = pb.PRIMES[-1]
    small = [p for p in range(2, 65537) if all(p % q for q in range(2, int(p ** 0.5) + 1))]

    def is_prime(x):
        return all(x % q for q in small if q * q <= x)
    expected = [x for x in range((1 << 32) - 1, lo - 1, -1) if is_prime(x)]
    assert pb.PRIMES == expected
    assert len(pb.PRIMES) == 48 and pb.PRIMES[0] == 4294967291
    assert all(a > b for a, b in itertools.pairwise(pb.PRIMES)) if hasattr(itertools, "pairwise") else True
