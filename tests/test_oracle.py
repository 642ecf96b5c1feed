"""Pins the CPU oracle (oracle/ww_oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that is not itself: the golden
fixtures (Python big-integer closed forms and the reference's own compiled CPU code,
tests/golden/make_golden.py), the known answers of the reference (tests.txt), published
prime counts, brute force on small inputs, and the compiled reference
(oracle/_ref) directly where it is built.
"""
import json
import os
import random
import sys

import pytest

import oracle

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
from make_golden import fib_mod  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "golden.json")))
SMALL = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79,
         83, 89, 97, 101, 103, 107, 109, 113]


def test_wieferich_and_wss_match_closed_forms():
    for row in GOLD["primes"]:
        p = row["p"]
        assert oracle.wieferich(p) == (row["s0"], row["s1"]), p
        assert oracle.wallsunsun(p) == (row["c20"], row["c21"]), p


def test_m2p_one_matches_closed_form():
    for row in GOLD["primes"]:
        assert oracle.m2p_one(row["p"]) == (row["one0"], row["one1"])


def test_checksum_slices_and_flags_match_reference_code():
    for row in GOLD["primes"]:
        fw, sw, fs, ss = oracle.cpu_flags(row["p"])
        assert (fw, sw, fs, ss) == (row["ref_wflag"], row["ref_wslice"], row["ref_sflag"],
                                    row["ref_sslice"]), row["p"]


def test_range_summaries_match_reference_code():
    for r in GOLD["ranges"]:
        prp, cks, nrec, flags = oracle.range_summary(r["lo"], r["hi"], threads=8)
        assert (prp, cks, flags) == (r["prp_count"], r["checksum"], r["cpu_result_count"])
        # for list members s0 is +-1, so every CPU flag is exactly one device record
        assert nrec == flags


def test_known_answer_special_instances():
    # tests.txt:16-18 and :23-25
    recs = oracle.range_results(2276306935816523 - 100, 2276306935816523 + 100)[2]
    assert (2276306935816523, 1093, -1, -3) in recs
    recs = oracle.range_results(338772621946054253, 338772621946054253 + 100)[2]
    assert (338772621946054253, 2, 0, 91) in recs
    # the two known Wieferich primes: 2^((p-1)/2) = +-1 mod p^2
    recs = oracle.range_results(1000, 4000)[2]
    assert (1093, 1093, 1, 0) in recs and (3511, 1093, 1, 0) in recs


def test_record_values_are_the_residue():
    """Every record states s = rem + quot*p (mod p^2) for Wieferich and
    F = quot*p (mod p^2) for Wall-Sun-Sun, with |quot| <= 1001. A Wieferich prime
    (s = +-1 mod p^2) is reported as rem 1, quot 0 for either sign
    (wieferich.cl:248-254)."""
    for lo, hi in [(127, 200_000), (10**12, 10**12 + 300_000)]:
        for p, code, rem, quot in oracle.range_results(lo, hi)[2]:
            p2 = p * p
            assert abs(quot) <= 1001
            if code == 1093:
                s = pow(2, (p - 1) // 2, p2)
                if (rem, quot) == (1, 0):
                    assert s in (1, p2 - 1)
                else:
                    assert s == (rem + quot * p) % p2
            else:
                kro = -1 if p % 5 in (2, 3) else 1
                f = fib_mod(p - kro, p2)
                assert f % p == 0 and f == (quot * p) % p2
                assert (code == 1) == (quot == 0)


def test_spsp2_membership():
    comps = GOLD["spsp2_composites_below_200001"]
    for n in range(3, 200_001, 2):
        want = n in comps or (all(n % q for q in range(3, int(n**0.5) + 1, 2)))
        assert oracle.strong_prp(2, n) == want, n
    # list members: no factor 3..113
    for n in (42799, 49141, 88357, 90751, 104653, 1373653, 25326001, 3215031751):
        assert oracle.is_prp(n) and not oracle.is_prime(n)
    for n in (2047, 3277, 4033, 4681, 8321):       # spsp(2) with a factor <= 113
        assert oracle.strong_prp(2, n) and not oracle.is_prp(n)
    for n in (49981, 83333, 129889):               # Fermat pseudoprimes, not strong
        assert pow(2, n - 1, n) == 1 and not oracle.is_prp(n)


def strong_prp_definition(b, n):
    """Miller-Rabin to base b from its definition, with Python's pow."""
    d, t = n - 1, 0
    while d % 2 == 0:
        d, t = d // 2, t + 1
    x = pow(b, d, n)
    if x in (1, n - 1):
        return True
    for _ in range(t - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def test_strong_prp_when_n_minus_1_is_a_power_of_two():
    """n = 2^t + 1, where the odd part of n - 1 is 1 (no exponent bits below the leading
    one): Fermat primes, 2^32 + 1 = 641 * 6700417 (a strong pseudoprime to base 2) and
    composites that are not."""
    for t in range(2, 64):
        n = (1 << t) + 1
        for b in (2, 3, 5):
            if n % b:
                assert oracle.strong_prp(b, n) == strong_prp_definition(b, n), (b, t)
    assert oracle.strong_prp(2, (1 << 32) + 1) and oracle.is_prp((1 << 32) + 1)
    assert not oracle.strong_prp(2, (1 << 40) + 1)


def test_ref_strong_prp_agrees_on_random_inputs():
    r = oracle.ref_lib()
    if r is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rng = random.Random(5)
    for _ in range(20000):
        n = rng.getrandbits(rng.randint(8, 64)) | 1
        if n < 3:
            continue
        assert bool(r.ref_strong_prp2(n)) == oracle.strong_prp(2, n), n


def test_is_prime_vs_trial_division_and_sieve():
    rng = random.Random(9)
    for _ in range(3000):
        n = rng.randrange(3, 1 << 40) | 1
        want = all(n % q for q in range(3, int(n**0.5) + 1, 2))
        assert oracle.is_prime(n) == want, n


def test_sieve_prime_counts():
    assert len(oracle.sieve_primes(0, 10**6)) == GOLD["pi"]["1e6"]
    assert len(oracle.sieve_primes(0, 10**7)) == GOLD["pi"]["1e7"]
    # brute force on windows far out
    for lo in (10**12, (1 << 40) + 12345):
        got = oracle.sieve_primes(lo, lo + 20000)
        want = [n for n in range(lo | 1, lo + 20000, 2) if oracle.is_prime(n)]
        assert got == want


def test_prime_count_127_to_1e8():
    """[127, 1e8): the list holds pi(1e8) - pi(126) = 5761425 primes plus the base-2
    strong pseudoprimes with no factor <= 113."""
    lst = oracle.prp_list(127, 10**8)
    primes = [p for p in lst if oracle.is_prime(p)]
    assert len(primes) == GOLD["pi"]["1e8"] - 30
    comps = [p for p in lst if not oracle.is_prime(p)]
    assert all(oracle.strong_prp(2, c) and all(c % q for q in SMALL) for c in comps)
    assert 42799 in comps and 25326001 in comps


def test_find_wheel_offset_brute_force():
    residues = [7, 11, 13, 17, 19, 23, 29, 1]
    for start in list(range(127, 3000)) + [10**12 + k for k in range(60)] + \
            [(1 << 63) + k for k in range(60)]:
        n, idx = oracle.find_wheel_offset(start)
        want = start
        while want % 2 == 0 or want % 3 == 0 or want % 5 == 0:
            want += 1
        assert n == want, start
        assert residues[idx] == n % 30, start


def test_candidate_predicate():
    for n in range(1, 5000):
        want = n % 2 == 1 and all(n % q for q in SMALL)
        assert oracle.is_candidate(n) == want
