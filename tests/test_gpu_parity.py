"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bit-exact everywhere: every quantity on this path is an unsigned integer. Small and
ragged ranges are compared entry by entry; the known-answer ranges of the reference
(tests.txt:6-25) and the full-size configurations are checked end to end.
"""
import json
import os
import random

import pytest

import oracle
import wwlib

GOLD = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                   "golden.json")))

pytestmark = pytest.mark.gpu

P64_MAX = 18446744073709551557  # largest prime below 2^64
T3_P = 2276306935816523
T4_P = 338772621946054253


def test_device_is_gfx950():
    info = wwlib.device_info(0)
    print("device:", info.name.decode(), info.arch.decode(), info.compute_units, "CUs")
    assert info.arch.decode().startswith("gfx950")


def _edge_prps():
    ps = [1093, 3511, 495791, 42799, 49141, 88357, 90751, 104653, 1373653, 25326001,
          3215031751, T3_P, T4_P, P64_MAX]
    for lo in (127, (1 << 32) - 4000, 1 << 32, (1 << 40) + 1, 1 << 50, (1 << 62) - 6000,
               1 << 63, (1 << 64) - 6000):
        ps += oracle.prp_list(lo, min(lo + 6000, (1 << 64) - 1))[:40]
    rng = random.Random(7)
    for bits in range(8, 65):
        for _ in range(6):
            lo = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
            lo = max(lo, 127)
            ps += oracle.prp_list(lo, min(lo + 4000, (1 << 64) - 1))[:3]
    return sorted(set(p for p in ps if p >= 127))


def test_residues_vs_oracle(gctx):
    ps = _edge_prps()
    got = gctx.residues(ps)
    bad = [(p, g, oracle.wieferich(p), oracle.wallsunsun(p)) for p, g in zip(ps, got)
           if g[:2] != oracle.wieferich(p) or g[2:] != oracle.wallsunsun(p)]
    assert not bad, bad[:5]
    print(f"{len(ps)} PRPs: residues bit-exact")


def test_prp_flags_vs_oracle(gctx):
    rng = random.Random(11)
    ns = [42799, 49141, 88357, 90751, 104653, 1373653, 25326001, 3215031751,  # spsp(2)
          49981, 83333, 129889,                                                # Fermat-only
          2047, 3277, 4033, 4681, 8321, 3, 5, 7, 9, 15, 127, P64_MAX, (1 << 64) - 1]
    for bits in range(3, 65):
        ns += [rng.getrandbits(bits) | (1 << (bits - 1)) | 1 for _ in range(200)]
    ns = [n for n in ns if n > 1]
    got = gctx.prp_flags(ns)
    want = [oracle.strong_prp(2, n) for n in ns]
    bad = [(n, g, w) for n, g, w in zip(ns, got, want) if g != w]
    assert not bad, bad[:10]


RANGES = [
    (127, 2_000_000),                          # 1093, 3511, every special instance p < 1001
    (127, 128), (127, 131), (128, 129), (1000, 1001), (1001, 1003),
    (2_000_000, 2_000_061), (5_000_001, 5_050_000),
    ((1 << 32) - (1 << 20), (1 << 32) + (1 << 20)),  # two-branch m2p_one boundary
    (1 << 50, (1 << 50) + (1 << 22)),          # 2^50 slice
    (T3_P - 1000, T3_P + 1000),                # Wieferich special instance (-1 -3 p)
    (T4_P, T4_P + 5000),                       # WSS special instance at START (inclusive)
    ((1 << 62) - 12345, (1 << 62) + 54321),
    (1 << 63, (1 << 63) + (1 << 22)),          # 2^63 slice
    ((1 << 63) - (1 << 20), (1 << 63) + (1 << 20)),  # 2^64 mod p: quotient 2 below, 1 above
    ((1 << 64) - (1 << 22), (1 << 64) - 1),    # top of the domain
    ((1 << 64) - 100, (1 << 64) - 1),
]


@pytest.mark.parametrize("lo,hi", RANGES)
def test_range_vs_oracle(gctx, lo, hi):
    got = gctx.run_range(lo, hi)
    prp, cks, recs = oracle.range_results(lo, hi)
    assert got.prp_count == prp
    assert got.checksum == cks, f"{got.checksum:016X} != {cks:016X}"
    assert got.results == recs


def test_ragged_ranges_vs_oracle(gctx):
    rng = random.Random(3)
    for _ in range(40):
        bits = rng.randint(8, 64)
        lo = max(127, rng.getrandbits(bits))
        w = rng.choice([1, 2, 3, 59, 60, 61, 127, 128, 129, 8191, 8192, 8193, 32767, 32768,
                        32769, 100_000])
        hi = min(lo + w, (1 << 64) - 1)
        if hi <= lo:
            continue
        got = gctx.run_range(lo, hi)
        prp, cks, recs = oracle.range_results(lo, hi)
        assert (got.prp_count, got.checksum, got.results) == (prp, cks, recs), (lo, hi)


@pytest.mark.parametrize("lo,hi", [(127, 3_000_000), (1 << 63, (1 << 63) + (1 << 21)),
                                   ((1 << 64) - (1 << 21), (1 << 64) - 1)])
def test_prp_list_vs_oracle_and_sieve(gctx, lo, hi):
    gctx.clear_results()
    gctx.clear_prp_count()
    gctx.getsegprimes(lo, hi)
    gctx.finish()
    got = sorted(gctx.prps())
    assert got == oracle.prp_list(lo, hi)
    # after CPU primality filtering the list is exactly the primes of the range
    # (checkPRPS, main.cpp:474-515, with a sieve of Eratosthenes for primesieve)
    primes = [p for p in got if wwlib.cpu_is_prime(p)]
    if hi < (1 << 40):
        assert primes == oracle.sieve_primes(lo, hi)
    else:
        assert primes == [p for p in got if oracle.is_prime(p)]


def test_prp_list_whole_range_127_1e8(gctx):
    """The whole [127, 10^8) range's PRP list (5,761,822 entries) from the sieve + strong PRP
    kernel equals the oracle's list, and after primality filtering it is exactly the
    sieve of Eratosthenes' primes of the range (pi(10^8) - 30 = 5,761,425); the rest are
    the range's base-2 strong pseudoprimes with no factor up to 113."""
    import numpy as np
    lo, hi = 127, 10**8
    gctx.clear_results()
    gctx.clear_prp_count()
    gctx.getsegprimes(lo, hi)
    gctx.finish()
    got = np.sort(np.array(gctx.prps(), dtype=np.uint64))
    want = np.array(oracle.prp_list(lo, hi), dtype=np.uint64)
    assert got.shape == want.shape and (got == want).all()
    primes = np.array(oracle.sieve_primes(lo, hi), dtype=np.uint64)
    assert len(primes) == 5_761_425
    assert np.isin(primes, got).all()
    rest = np.setdiff1d(got, primes)
    assert len(got) == len(primes) + len(rest)
    assert all(oracle.is_prp(int(n)) and not oracle.is_prime(int(n)) for n in rest)


def test_chunking_invariance(gctx):
    lo, hi = 10**13, 10**13 + 3 * (1 << 22) + 12345
    ref = gctx.run_range(lo, hi)
    for chunk in (1 << 20, 999_983, 1 << 23):
        got = gctx.run_range(lo, hi, chunk)
        assert (got.prp_count, got.checksum, got.results) == (ref.prp_count, ref.checksum,
                                                              ref.results)


# Known answers of the reference (tests.txt): START END -> results file lines.
KAT = [
    ("T1", 18446744013709551615, 18446744073709551615, [], 0x509DE6503B2A6E8E),
    ("T2", 127, 60_000_000_000, None, 0x8B77D2DA053842DD),  # START 1 clamps to 127
    ("T3", 2276300000000000, 2276360000000000, [(T3_P, 1093, -1, -3)], 0x6522CBC4CA2E4CFF),
    ("T4", 338772621946054253, 338772681946054253, [(T4_P, 2, 0, 91)], 0x589A69274A6AB5A8),
]


@pytest.mark.parametrize("name,lo,hi,lines,cks", KAT, ids=[k[0] for k in KAT])
def test_known_answers(gctx, name, lo, hi, lines, cks):
    got = gctx.run_range(lo, hi, 1 << 32)
    print(f"{name}: {got.prp_count} PRPs, checksum {got.checksum:016X}")
    assert got.checksum == cks, f"{got.checksum:016X} != {cks:016X}"
    primes_only = [r for r in got.results if wwlib.cpu_is_prime(r[0])]
    if lines is not None:
        assert primes_only == lines


def test_2p33_range_at_2p63_properties(gctx):
    """[2^63, 2^63 + 2^33): checksum additivity over halves, and 16 sampled 2^22-wide
    slices bit-exact against the oracle."""
    lo, hi = 1 << 63, (1 << 63) + (1 << 33)
    whole = gctx.run_range(lo, hi, 1 << 32)
    mid = lo + (1 << 32) + 777
    a = gctx.run_range(lo, mid, 1 << 32)
    b = gctx.run_range(mid, hi, 1 << 32)
    assert whole.prp_count == a.prp_count + b.prp_count
    assert whole.checksum == (a.checksum + b.checksum) % (1 << 64)
    print(f"[2^63, 2^63+2^33): {whole.prp_count} PRPs, checksum {whole.checksum:016X}")
    for k in range(16):
        s = lo + k * (1 << 29)
        got = gctx.run_range(s, s + (1 << 22))
        prp, cks, _, _ = oracle.range_summary(s, s + (1 << 22), threads=16)
        assert (got.prp_count, got.checksum) == (prp, cks), k


@pytest.mark.parametrize("whole", GOLD["full_ranges"], ids=lambda c: c["name"])
def test_whole_range_vs_reference_code(gctx, whole):
    """[127, 10^8) and [2^50, 2^50 + 2^32) whole, plain and persistent forms, against the
    reference's own CPU path (compiled from wwcpu.cpp / prime.h, tests/golden)."""
    lo, hi = whole["lo"], whole["hi"]
    got = gctx.run_range(lo, hi, 1 << 32)
    print(f"{whole['name']}: {got.prp_count} PRPs, checksum {got.checksum:016X}, "
          f"{len(got.results)} records")
    assert (got.prp_count, got.checksum) == (whole["prp_count"], whole["checksum"])
    assert len(got.results) == whole["cpu_result_count"]
    with wwlib.Batch(gctx, [(lo, hi)]) as b:
        b.launch(graph=True)
        (prp, cks, recs), = b.read()
    assert (prp, cks) == (whole["prp_count"], whole["checksum"])
    assert recs == got.results


def test_fused_chunks_vs_oracle(gctx):
    """ww_search_fused: the persistent kernel as the body of the chunk loop adds the same
    counters and records to the context as the two-kernel form, over the edge ranges, in
    uneven chunks, and mixed with two-kernel chunks between clears."""
    for lo, hi in RANGES:
        got = gctx.run_range_fused(lo, hi)
        assert (got.prp_count, got.checksum, got.results) == oracle.range_results(lo, hi), (lo, hi)
    lo, hi = (1 << 40) + 17, (1 << 40) + 17 + 3_000_000
    want = oracle.range_results(lo, hi)
    got = gctx.run_range_fused(lo, hi, chunk=1_000_003)
    assert (got.prp_count, got.checksum, got.results) == want
    gctx.clear_results()
    gctx.clear_prp_count()
    gctx.search_fused(lo, lo + 1_500_000)
    gctx.search_range(lo + 1_500_000, hi, 1 << 32)
    gctx.finish()
    c = gctx.counters()
    assert (c.totalprimecount, c.checksum) == want[:2]
    assert wwlib.sort_results(gctx.results()) == want[2]


def test_fused_chunks_at_full_size(gctx):
    lo, hi = 1 << 63, (1 << 63) + (1 << 33)
    fused = gctx.run_range_fused(lo, hi)
    plain = gctx.run_range(lo, hi, 1 << 32)
    assert (fused.prp_count, fused.checksum, fused.results) == \
        (plain.prp_count, plain.checksum, plain.results)
    assert f"{fused.checksum:016X}" == "0BB982E18153E292"


def test_random_ranges_all_forms_vs_oracle(gctx):
    """200 random ranges across the 64-bit domain (widths 1 to 2^20 above 2^20, up to 2^12
    below it, where nearly every prime yields records): the two-kernel form, fused chunks
    and one persistent batch all give the oracle's PRP count, checksum and records."""
    rng = random.Random(31)
    ranges = []
    while len(ranges) < 200:
        bits = rng.randint(7, 64)
        lo = max(127, rng.getrandbits(bits))
        width = rng.randint(1, 1 << (20 if bits > 20 else 12))
        hi = min(lo + width, (1 << 64) - 1)
        if hi > lo:
            ranges.append((lo, hi))
    with wwlib.Batch(gctx, ranges) as b:
        b.launch(graph=True)
        batched = b.read()
    for (lo, hi), bat in zip(ranges, batched):
        want = oracle.range_results(lo, hi)
        plain = gctx.run_range(lo, hi)
        fused = gctx.run_range_fused(lo, hi)
        assert (plain.prp_count, plain.checksum, plain.results) == want, (lo, hi)
        assert (fused.prp_count, fused.checksum, fused.results) == want, (lo, hi)
        assert bat == want, (lo, hi)


@pytest.mark.parametrize("top", [62, 40, 20])
def test_wss_table_start_switch_vs_oracle(gctx, top):
    """Around the p where the six leading bits of the Wall-Sun-Sun ladder's exponent
    n = (p - (p/5)) / 2 go from k = 45 (six table steps allowed) to k = 46 (five only):
    waves there hold lanes of both kinds and take five steps for all of them. Residues,
    the two-kernel form and a persistent batch against the oracle."""
    p_switch = 46 << (top - 4)  # n = 46 * 2^(top-5), p = 2n +- 1
    lo, hi = p_switch - 20000, p_switch + 20000
    ps = oracle.prp_list(lo, hi)
    ks = {(((p + 1 if p % 5 in (2, 3) else p - 1) // 2) >> (top - 5)) for p in ps}
    assert {45, 46} <= ks, ks
    got = gctx.residues(ps)
    for p, g in zip(ps, got):
        assert g == tuple(oracle.wieferich(p)) + tuple(oracle.wallsunsun(p)), p
    want = oracle.range_results(lo, hi)
    plain = gctx.run_range(lo, hi)
    assert (plain.prp_count, plain.checksum, plain.results) == want
    with wwlib.Batch(gctx, [(lo, hi)]) as b:
        b.launch()
        assert b.read() == [want]
