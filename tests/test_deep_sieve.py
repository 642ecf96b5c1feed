"""The deep sieve's rescue classes (ww_deep_table, host side) and its exactness: a
sieve by 127..B that clears each prime's base-2 pseudoprime class again never removes a
list member. The oracle's PRP lists are the truth; big-integer arithmetic checks the
table entries."""
import random

import oracle
import wwlib

KNOWN_SPSP2 = [42799, 49141, 88357, 90751, 104653, 130561, 196093,
               1373653, 25326001, 3215031751, 4294967297, 2152302898747, 3474749660383]


def order2(ell):
    k, x = 1, 2 % ell
    while x != 1:
        x = x * 2 % ell
        k += 1
    return k


def test_table_entries():
    t = wwlib.deep_table(8192)
    assert len(t) == 1028 - 30  # pi(8192) - pi(113)
    for ell, m, rho, c32m, bar_ell, bar_m in t:
        o = order2(ell)
        lcm = o if o % 2 == 0 else 2 * o
        assert m == ell * lcm and 0 <= rho < m
        assert rho % ell == 0 and rho % lcm == 1          # the class of the pseudoprimes
        assert rho == ell
        assert pow(2, rho - 1, ell) == 1
        assert c32m == (1 << 32) % m
        assert bar_ell == (1 << 32) // ell and bar_m == (1 << 64) // m
    assert [e for e, *_ in wwlib.deep_table(4096)] == [e for e, *_ in t if e <= 4096]


def test_table_to_the_largest_bound():
    t = wwlib.deep_table(65535)
    assert len(t) == 6542 - 30  # pi(2^16) - pi(113)
    assert t[-1][0] == 65521
    rng = random.Random(5)
    for ell, m, rho, c32m, bar_ell, bar_m in rng.sample(t, 200) + t[-5:]:
        o = order2(ell)
        assert m == ell * (o if o % 2 == 0 else 2 * o) and m < (1 << 32)
        assert rho == ell and c32m == (1 << 32) % m
        assert bar_ell == (1 << 32) // ell and bar_m == (1 << 64) // m


def order2_factored(ell):
    """ord_ell(2) from the prime factors of ell - 1 (for primes too large to step)."""
    o, n, q = ell - 1, ell - 1, 2
    while q * q <= n:
        if n % q == 0:
            while n % q == 0:
                n //= q
            while o % q == 0 and pow(2, o // q, ell) == 1:
                o //= q
        q += 1
    if n > 1 and pow(2, o // n, ell) == 1:
        o //= n
    return o


def test_presieve_table_entries():
    """ww_presieve_table: every prime in (above, bound], with m = ell * lcm(2, ord_ell(2))
    and the 64-bit Barrett reciprocals of ell and m."""
    t = wwlib.presieve_table(2048, 1 << 20)
    assert len(t) == 82025 - 309  # pi(2^20) - pi(2048)
    assert t[0][0] == 2053 and t[-1][0] == 1048573
    assert all(a[0] < b[0] for a, b in zip(t, t[1:]))
    rng = random.Random(6)
    for ell, m, bar_ell, bar_m in t[:40] + rng.sample(t, 400) + t[-5:]:
        assert oracle.is_prime(ell)
        o = order2_factored(ell)
        if ell < 20000:
            assert o == order2(ell)
        assert pow(2, o, ell) == 1
        assert m == ell * (o if o % 2 == 0 else 2 * o)
        assert bar_ell == (1 << 64) // ell and bar_m == (1 << 64) // m
    top = wwlib.presieve_table(1 << 20, 1 << 22)
    assert len(top) == 295947 - 82025 and top[-1][0] == 4194301  # pi(2^22) - pi(2^20)
    assert [e for e, *_ in wwlib.presieve_table(2048, 4096)] == [e for e, *_ in t if e <= 4096]


def survivors(lo, hi, bound):
    """Odd n in [lo, hi) with no factor 3..113 left by the deep sieve to `bound`
    (every multiple of a deep prime marked, then its class rho mod m cleared)."""
    t = [(e, m, r) for e, m, r, *_ in wwlib.deep_table(bound)]
    base = lo | 1
    keep = {}
    for n in range(base, hi, 2):
        keep[n] = oracle.is_candidate(n)
    marked = set()
    for ell, m, rho in t:
        first = -(-base // ell) * ell
        if first % 2 == 0:
            first += ell
        for n in range(first, hi, 2 * ell):
            if n != ell:
                marked.add(n)
    for ell, m, rho in t:
        first = base + ((rho - base) % m)
        for n in range(first, hi, m):
            marked.discard(n)
    return [n for n in range(base, hi, 2) if keep[n] and n not in marked]


def test_known_pseudoprimes_survive():
    for n in KNOWN_SPSP2:
        assert oracle.is_prp(n) and not oracle.is_prime(n)
        assert n in survivors(n - 300, n + 300, 8192), n


def test_survivors_contain_the_list_and_thin_candidates():
    rng = random.Random(4)
    wins = [(127, 60000)] + [(w, w + 20000) for w in
                             (rng.randrange(1 << 20, 1 << 62) for _ in range(6))]
    for lo, hi in wins:
        lst = oracle.prp_list(lo, hi)
        cand = sum(oracle.is_candidate(n) for n in range(lo | 1, hi, 2))
        for bound in (127, 1024, 4096, 8192):
            s = survivors(lo, hi, bound)
            assert set(lst) <= set(s), (lo, hi, bound)
        if lo > 10**9:
            # Mertens: prod (1 - 1/l) over 113 < l <= 4096 ~ 0.57 of the candidates remain
            assert len(survivors(lo, hi, 4096)) < 0.65 * cand
