"""The bucket sieve on the device (ww_bucket.hip, the pre-sieve's apply, k_batch's known
primes): the table of odd primes below 2^32, and batch results equal to the oracle's
wherever the bucket sieve runs, including ranges that hold base-2 strong pseudoprimes whose
factors lie in each sieve tier (which the tiers must keep for the strong PRP test), the
plain form's PRP list and fused chunks with the bucket sieve, mixed batches, the
configurations that turn it off, graph replays, and overflowing buckets."""
import random

import numpy as np
import pytest

import oracle
import wwlib

pytestmark = pytest.mark.gpu


@pytest.fixture
def bctx(gctx):
    """The session context with the bucket sieve on for ranges of any width."""
    saved = (gctx.bucket_sieve(), gctx.deep_bound(), gctx.presieve_bound(), gctx.launch_shape())
    gctx.set_bucket_sieve(0)
    yield gctx
    gctx.set_bucket_sieve(saved[0])
    gctx.set_bucket_capacity(1.0)
    gctx.set_deep_bound(saved[1])
    gctx.set_presieve_bound(saved[2])
    gctx.set_launch_shape(*saved[3])


def odd_primes(lo, hi):
    """Odd primes in [lo, hi) (hi <= 2^32) by a numpy segmented sieve."""
    lim = int(hi ** 0.5) + 1
    small = np.ones(lim + 1, dtype=bool)
    small[:2] = False
    for i in range(2, int(lim ** 0.5) + 1):
        if small[i]:
            small[i * i::i] = False
    seg = np.ones(hi - lo, dtype=bool)
    for q in np.nonzero(small)[0]:
        q = int(q)
        start = max(q * q, (lo + q - 1) // q * q)
        seg[start - lo::q] = False
    if lo <= 1:
        seg[:2 - lo] = False
    vals = np.nonzero(seg)[0] + lo
    return [int(v) for v in vals if v % 2]


def test_prime_table(gctx):
    head, total = gctx.prime_table(0, 10)
    assert total == 203280221 - 1  # pi(2^32), without 2
    assert head == [3, 5, 7, 11, 13, 17, 19, 23, 29, 31]
    tail, _ = gctx.prime_table(total - 3, 3)
    assert tail == [4294967231, 4294967279, 4294967291]
    # the primes around the pre-sieve bound 2^20 (index 82025 - 1), and random slices
    # against a numpy sieve of the same values
    first = 82025 - 1
    got, _ = gctx.prime_table(first - 1000, 3000)
    assert got == odd_primes(got[0], got[-1] + 1)
    assert got[999] == 1048573 and got[1000] == 1048583  # last below 2^20, first above
    rng = random.Random(3)
    for _ in range(6):
        at = rng.randrange(0, total - 20000)
        got, _ = gctx.prime_table(at, 20000)
        assert got == odd_primes(got[0], got[-1] + 1), at


def test_order_table(gctx):
    """The device's order table against the host's order arithmetic (itself checked against
    the definition on the CPU): zeros up to 2^20, then slices at the tier's start, at random
    places and at the end of the table."""
    _, total = gctx.prime_table(0, 1)
    first = 82025 - 1  # first table prime above 2^20
    got, n = gctx.prime_orders(first - 5, 4005)
    assert n == total and got[:5] == [0] * 5
    ps, _ = gctx.prime_table(first, 4000)
    assert ps[0] == 1048583
    assert got[5:] == [wwlib.cpu_order2(l) for l in ps]
    rng = random.Random(4)
    for at in [rng.randrange(first, total - 3000) for _ in range(5)] + [total - 3000]:
        ps, _ = gctx.prime_table(at, 3000)
        got, _ = gctx.prime_orders(at, 3000)
        assert got == [wwlib.cpu_order2(l) for l in ps], at
    # primes of small order: 2^31 - 1 and 6700417 (a factor of 2^32 + 1), found in the
    # table by bisection
    for ell, order in ((2147483647, 31), (6700417, 64)):
        a, b = 0, total
        while b - a > 1:
            mid = (a + b) // 2
            if gctx.prime_table(mid, 1)[0][0] <= ell:
                a = mid
            else:
                b = mid
        assert gctx.prime_table(a, 1)[0] == [ell]
        assert gctx.prime_orders(a, 1)[0] == [order]


RANGES = [
    (10**12, 10**12 + (1 << 22)),                              # top 10^6: pre-sieve tier only
    (1 << 63, (1 << 63) + (1 << 23)),                          # sqrt ~ 3.04e9
    ((1 << 64) - (1 << 23) - 1, (1 << 64) - 1),                # sqrt ~ 2^32: every table prime
    (127, 1 << 22),                                            # the domain's start
    ((1 << 40) + 12345, (1 << 40) + 12345 + (1 << 23)),       # top just above 2^20
    ((1 << 50) + 7, (1 << 50) + 7 + 3 * (1 << 21) + 4097),    # ragged end, short last piece
]


@pytest.mark.parametrize("lo,hi", RANGES)
def test_bucket_ranges_vs_oracle(bctx, lo, hi):
    with wwlib.Batch(bctx, [(lo, hi)]) as b:
        b.launch()
        got = b.read()[0]
        info = b.buckets()
    assert info["spans"] == 1 and info["active"] and info["overflows"] == 0, info
    assert got == oracle.range_results(lo, hi), (lo, hi)


def chernick_spsp2():
    """(6k+1)(12k+1)(18k+1), all three prime (above 2^20) and a base-2 strong pseudoprime."""
    out = []
    for k in range(174763, 200000):
        a, b, c = 6 * k + 1, 12 * k + 1, 18 * k + 1
        if oracle.is_prime(a) and oracle.is_prime(b) and oracle.is_prime(c):
            n = a * b * c
            if oracle.strong_prp(2, n):
                out.append(n)
        if len(out) == 3:
            return out
    return out


def twin_spsp2(p0, count):
    """p (2p - 1), both prime, a base-2 strong pseudoprime: the smaller factor p from p0 on."""
    out, p = [], p0 | 1
    while len(out) < count:
        if oracle.is_prime(p) and oracle.is_prime(2 * p - 1):
            n = p * (2 * p - 1)
            if n < 1 << 64 and oracle.strong_prp(2, n):
                out.append((p, n))
        p += 2
    return out


def test_strong_pseudoprimes_in_every_tier_are_kept(bctx):
    """Ranges around base-2 strong pseudoprimes whose smallest factor lies in the deep tier
    (127..2048), the pre-sieve tier (2048..2^20) or the bucket tier (2^20..2^32): each must
    be in the list (the oracle's list holds it), which the count and checksum show."""
    cases = [n for _, n in twin_spsp2(1500, 2)]           # deep tier: 1657, 2089
    cases += [n for _, n in twin_spsp2(70001, 2)]         # pre-sieve tier
    cases += [n for _, n in twin_spsp2(2100000, 2)]       # bucket tier, n ~ 2^43
    cases += [n for _, n in twin_spsp2(2500000000, 2)]    # bucket tier, n ~ 1.25e19
    cases += [n for _, n in twin_spsp2(3027000000, 1)]    # bucket tier, p near sqrt(2^64)
    cases += chernick_spsp2()                             # three factors above 2^20
    cases += [3215031751, 2152302898747, 3474749660383, 4294967297]
    ranges = [(max(127, n - (1 << 21)), n + (1 << 21)) for n in cases]
    with wwlib.Batch(bctx, ranges) as b:
        b.launch()
        got = b.read()
        info = b.buckets()
    assert info["active"] and info["spans"] == len(ranges) and info["overflows"] == 0
    for n, (lo, hi), g in zip(cases, ranges, got):
        assert not oracle.is_prime(n) and oracle.strong_prp(2, n)
        assert n in oracle.prp_list(lo, hi)
        assert g == oracle.range_results(lo, hi), n


@pytest.mark.parametrize("lo,hi", RANGES)
def test_fused_chunks_with_buckets_vs_oracle(bctx, lo, hi):
    """ww_search_fused (the drop-in's chunk body) with the bucket sieve: the same counters
    and records as the oracle, in one chunk and in uneven ones."""
    want = oracle.range_results(lo, hi)
    got = bctx.run_range_fused(lo, hi)
    assert (got.prp_count, got.checksum, got.results) == want, (lo, hi)
    got = bctx.run_range_fused(lo, hi, chunk=(hi - lo) // 3 + 12345)
    assert (got.prp_count, got.checksum, got.results) == want, (lo, hi)


def test_fused_chunks_keep_strong_pseudoprimes(bctx):
    cases = [n for _, n in twin_spsp2(1500, 1)] + [n for _, n in twin_spsp2(2100000, 1)]
    cases += [n for _, n in twin_spsp2(2500000000, 1)] + chernick_spsp2()[:1]
    for n in cases:
        lo, hi = n - (1 << 21), n + (1 << 21)
        got = bctx.run_range_fused(lo, hi)
        assert (got.prp_count, got.checksum, got.results) == oracle.range_results(lo, hi), n


def test_fused_overflowing_buckets_stay_exact(bctx):
    bctx.set_bucket_capacity(0.1)
    lo, hi = (1 << 63) + 1001, (1 << 63) + 1001 + (1 << 23)
    got = bctx.run_range_fused(lo, hi)
    assert (got.prp_count, got.checksum, got.results) == oracle.range_results(lo, hi)


def plain_list(ctx, lo, hi):
    """ww_getsegprimes over [lo, hi) into an empty list, read back sorted."""
    ctx.clear_prp_count()
    ctx.getsegprimes(lo, hi)
    ctx.finish()
    return sorted(ctx.prps())


@pytest.mark.parametrize("lo,hi", RANGES)
def test_plain_form_with_buckets_vs_oracle(bctx, lo, hi):
    """ww_getsegprimes with the bucket sieve: known primes join the list without the strong
    PRP test and flagged candidates are tested, so the list is the oracle's; the two-kernel
    form's counters and records equal the oracle's, in one chunk and in uneven ones."""
    assert plain_list(bctx, lo, hi) == oracle.prp_list(lo, hi), (lo, hi)
    want = oracle.range_results(lo, hi)
    for chunk in (hi - lo, (hi - lo) // 3 + 12345):
        got = bctx.run_range(lo, hi, chunk)
        assert (got.prp_count, got.checksum, got.results) == want, (lo, hi, chunk)


def test_plain_form_keeps_strong_pseudoprimes(bctx):
    cases = [n for _, n in twin_spsp2(1500, 1)] + [n for _, n in twin_spsp2(70001, 1)]
    cases += [n for _, n in twin_spsp2(2100000, 1)] + [n for _, n in twin_spsp2(2500000000, 1)]
    cases += chernick_spsp2()[:1]
    for n in cases:
        lo, hi = n - (1 << 21), n + (1 << 21)
        got = plain_list(bctx, lo, hi)
        assert n in got and got == oracle.prp_list(lo, hi), n


def test_plain_form_overflowing_buckets_stay_exact(bctx):
    bctx.set_bucket_capacity(0.1)
    lo, hi = (1 << 63) + 1001, (1 << 63) + 1001 + (1 << 23)
    assert plain_list(bctx, lo, hi) == oracle.prp_list(lo, hi)
    got = bctx.run_range(lo, hi)
    assert (got.prp_count, got.checksum, got.results) == oracle.range_results(lo, hi)


def test_mixed_batch_vs_oracle(bctx):
    """Bucket spans beside ranges too narrow for them, some sharing a pre-sieve chunk."""
    bctx.set_bucket_sieve(1 << 22)
    rng = random.Random(11)
    ranges = [((1 << 63) + 5, (1 << 63) + 5 + (1 << 22) + 8191),
              (10**15, 10**15 + 30000),
              ((1 << 62) + 1, (1 << 62) + 1 + (1 << 22)),
              ((1 << 62) + (1 << 22) + 1, (1 << 62) + (1 << 23)),
              (2**33 + 1, 2**33 + 1 + (1 << 23)),
              (997, 5000)]
    for _ in range(8):
        lo = rng.getrandbits(63) | 1
        ranges.append((lo, lo + rng.choice([100, 70000, 1 << 22, (1 << 22) + 12345])))
    with wwlib.Batch(bctx, ranges) as b:
        b.launch()
        got = b.read()
        info = b.buckets()
    assert info["active"] and info["spans"] == sum(hi - lo >= 1 << 22 for lo, hi in ranges)
    for (lo, hi), g in zip(ranges, got):
        assert g == oracle.range_results(lo, hi), (lo, hi)


@pytest.mark.parametrize("setting", ["deep0", "deep512", "pre19", "seg19"])
def test_configurations_without_buckets(bctx, setting):
    """Sieve tiers that leave a gap below sqrt(hi) (deep bound under 1024, pre-sieve bound
    under 2^20) or a pre-sieve segment other than 2^20 bits: the batch runs without known
    primes, with the same results."""
    lo, hi = (1 << 63) + 99, (1 << 63) + 99 + (1 << 22)
    if setting == "deep0":
        bctx.set_deep_bound(0)
    elif setting == "deep512":
        bctx.set_deep_bound(512)
    elif setting == "pre19":
        bctx.set_presieve_bound(1 << 19)
    else:
        bctx.set_launch_shape(4, 19)
    with wwlib.Batch(bctx, [(lo, hi)]) as b:
        b.launch()
        got = b.read()[0]
        info = b.buckets()
    assert info["spans"] == 1 and not info["active"]
    assert got == oracle.range_results(lo, hi)


def test_graph_replays_equal_eager(bctx):
    ranges = [((1 << 61) + 3, (1 << 61) + 3 + (1 << 22)), (10**14, 10**14 + (1 << 21) + 5)]
    with wwlib.Batch(bctx, ranges) as b:
        b.launch()
        eager = b.read()
        for _ in range(3):
            b.launch(graph=True)
            assert b.read() == eager
            bctx.run_range(10**12, 10**12 + 12345)  # eager work between replays
    assert eager == [oracle.range_results(lo, hi) for lo, hi in ranges]


def test_overflowing_buckets_stay_exact(bctx):
    """Buckets at a tenth of their size overflow; the chunks concerned test every candidate
    and the results do not change."""
    bctx.set_bucket_capacity(0.1)
    lo, hi = (1 << 63) + 1001, (1 << 63) + 1001 + (1 << 23)
    with wwlib.Batch(bctx, [(lo, hi)]) as b:
        b.launch()
        got = b.read()[0]
        info = b.buckets()
    assert info["active"] and info["overflows"] > 0, info
    assert got == oracle.range_results(lo, hi)
