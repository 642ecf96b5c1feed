"""The bucket sieve's walk of one sieving prime (csrc/ww_bucket.h), built for the host
through ww_cpu_bucket_hits: for a prime l above 2^20 and the odd numbers g0 + 2b, it must
visit exactly the multiples n = l*k with k prime to 30030 x 17 x 19 x ... x 53 (in order,
as bit indices (n - g0) / 2), and flag exactly those with 2^(n-1) = 1 (mod l), the prime's
base-2 pseudoprime class. The expected lists come from Python's exact integers: every
multiple of l in the window, a gcd with 30030 x 17 x 19 x ... x 53, and pow(2, n - 1, l).
Primes of small multiplicative order (factors of 2^31 - 1, 2^37 - 1, 2^41 - 1, ...) make the
class frequent, windows near 2^64 make the multiples pass 2^64. The walk's class test uses the prime's base-2 order, which
ww_cpu_order2 computes with the order table's arithmetic; it is checked against the
definition of the order. No device needed."""
import math
import random

import pytest

import oracle
import wwlib


def brute(g0, nbits, ell):
    out = []
    k = -(-g0 // ell)
    while ell * k < g0 + 2 * nbits:
        if math.gcd(k, 30030 * 17 * 19 * 23 * 29 * 31 * 37 * 41 * 43 * 47 * 53) == 1:
            n = ell * k
            out.append(((n - g0) // 2, pow(2, n - 1, ell) == 1))
        k += 1
    return out


# primes above 2^20 with 2 of small order: 2^31 - 1 (31), 616318177 | 2^37 - 1 (37),
# 164511353 | 2^41 - 1 (41), 2099863 | 2^43 - 1 (43), 13264529 | 2^47 - 1 (47),
# 20394401 | 2^53 - 1 (53), 6700417 | 2^32 + 1 (64)
SMALL_ORDER = [2147483647, 616318177, 164511353, 2099863, 13264529, 20394401, 6700417]


def test_small_order_primes_flag_their_class():
    for ell in SMALL_ORDER:
        order = next(d for d in range(1, 65) if pow(2, d, ell) == 1)
        assert order <= 64
        for g0 in (1, (1 << 20) + 1, (1 << 63) + 1, (1 << 64) - (1 << 30) + 1, 10**15 + 1,
                   ((1 << 64) - 700 * ell) | 1):
            got = wwlib.cpu_bucket_hits(g0, 700 * ell, ell, cap=1 << 16)
            assert got == brute(g0, 700 * ell, ell), (ell, g0)
            assert any(r for _, r in got), (ell, g0)  # ~1 in 11 hits is in the class


def test_random_primes_and_windows():
    rng = random.Random(5)

    def rprime(lo, hi):
        while True:
            x = rng.randrange(lo, hi) | 1
            if oracle.is_prime(x):
                return x

    for _ in range(60):
        ell = rprime((1 << 20) + 1, 1 << 32)
        g0 = rng.choice([rng.randrange(1, 1 << 64) | 1,
                         ((1 << 64) - rng.randrange(1, 1 << 34)) | 1,
                         rng.randrange(1, 1 << 40) | 1])
        nbits = rng.choice([1, 100, ell * 20, ell * 300])
        assert wwlib.cpu_bucket_hits(g0, nbits, ell, cap=1 << 16) == brute(g0, nbits, ell), \
            (g0, nbits, ell)


def test_extreme_primes():
    # the largest prime below 2^32 over the top of the domain, the smallest above 2^20
    g0 = ((1 << 64) - (1 << 33)) | 1
    assert wwlib.cpu_bucket_hits(g0, 1 << 32, 4294967291) == brute(g0, 1 << 32, 4294967291)
    g0 = (1 << 63) + 1
    got = wwlib.cpu_bucket_hits(g0, 1 << 32, 1048583, cap=1 << 16)
    assert got == brute(g0, 1 << 32, 1048583)
    assert len(got) > 1050  # 2^32 / l odd multiples, 0.272 of them visited


def prime_factors(n):
    out, q = [], 2
    while q * q <= n:
        if n % q == 0:
            out.append(q)
            while n % q == 0:
                n //= q
        q += 1 if q == 2 else 2
    if n > 1:
        out.append(n)
    return out


def check_order(ell, o):
    """o is ord_ell(2) by the definition: 2^o = 1, and 2^(o/q) != 1 for every prime q | o."""
    assert (ell - 1) % o == 0 and pow(2, o, ell) == 1, (ell, o)
    for q in prime_factors(o):
        assert pow(2, o // q, ell) != 1, (ell, o, q)


def test_order_small_and_structured():
    # small orders, found by direct enumeration
    for ell in SMALL_ORDER:
        o = wwlib.cpu_order2(ell)
        assert o == next(d for d in range(1, 65) if pow(2, d, ell) == 1)
    # l - 1 with a large power of two (3 2^30 + 1, 15 2^27 + 1, 7 2^26 + 1, 5 2^25 + 1,
    # 119 2^23 + 1), safe primes (l - 1 = 2 x prime), the extremes of the bucket tier
    cases = [3221225473, 2013265921, 469762049, 167772161, 998244353,
             4294967291, 1048583, 2147483647, 4294967279]
    p = (1 << 31) + 3  # 11 mod 12, as every safe prime above 5
    while len(cases) < 20:  # safe primes above 2^31
        if oracle.is_prime(p) and oracle.is_prime((p - 1) // 2):
            cases.append(p)
        p += 12
    for ell in cases:
        assert oracle.is_prime(ell)
        check_order(ell, wwlib.cpu_order2(ell))


def test_order_repeated_factors():
    """Primes whose order drops a factor 3 or 5 more than once (q^2 | (l - 1) / ord), so
    the reduction must repeat; and primes where 2 is a quadratic residue (l = +-1 mod 8),
    which halve the order's power of two."""
    rng = random.Random(7)
    found = {3: 0, 5: 0, 2: 0}
    while min(found.values()) < 4:
        ell = rng.randrange((1 << 20) + 1, 1 << 32) | 1
        if not oracle.is_prime(ell):
            continue
        o = wwlib.cpu_order2(ell)
        check_order(ell, o)
        for q in (3, 5):
            if (ell - 1) % (q * q) == 0 and ((ell - 1) // o) % (q * q) == 0:
                found[q] += 1
        if ell % 8 in (1, 7) and ((ell - 1) // o) % 4 == 0:
            found[2] += 1


def test_order_random():
    rng = random.Random(9)
    for _ in range(400):
        ell = rng.randrange((1 << 20) + 1, 1 << 32) | 1
        while not oracle.is_prime(ell):
            ell += 2
        if ell >= 1 << 32:
            continue
        check_order(ell, wwlib.cpu_order2(ell))


def test_longest_window():
    """A window of 30029 cofactors, the most a walk may span, starting at the wheel's last
    residue so that the walk's index runs through both copies of the gaps; one more bit is
    refused."""
    ell = 1048583
    nbits = 30028 * ell // 2
    for k0 in (30029, 30030 * 7 + 30029, 30030 * 1000 + 1):
        g0 = ell * k0
        got = wwlib.cpu_bucket_hits(g0, nbits, ell, cap=1 << 16)
        assert got == brute(g0, nbits, ell), k0
        assert len(got) > 3900  # about 5760 x 16/17 x 18/19 x ... x 52/53
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_bucket_hits(ell * 30029, nbits + 1, ell)


def test_arguments():
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_bucket_hits(2, 100, 2147483647)  # even start
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_bucket_hits(1, 100, 1 << 19)  # below the bucket tier
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_order2(1 << 19)
