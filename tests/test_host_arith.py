"""The device arithmetic (csrc/ww_arith.h) built for the host, through ww_cpu_residues and
ww_cpu_spsp2: the Wieferich and Wall-Sun-Sun ladders and the windowed strong PRP test
give the oracle's values from every leading-bit bound a wave may share (any value from
the lane's own up to 63), including bounds that cross the 32-bit word boundary the
Wall-Sun-Sun ladder reads its bits by, and with either length of the Wall-Sun-Sun
ladder's table start. No device needed."""
import random

import pytest

import oracle
import wwlib


def sample_primes(rng, count):
    ps = []
    while len(ps) < count:
        bits = rng.choice([8, 12, 20, 31, 32, 33, 40, 50, 62, 63, 64])
        p = rng.getrandbits(bits) | (1 << (bits - 1)) | 1
        if p >= 127 and p % 5 and oracle.is_prp(p):
            ps.append(p)
    return ps


def lead(x):
    return x.bit_length() - 1


def wss_lead(p):
    n = (p + 1 if p % 5 in (2, 3) else p - 1) // 2
    return lead(n)


def test_ladders_match_oracle_at_own_bounds():
    rng = random.Random(21)
    ps = sample_primes(rng, 600) + [1093, 3511, (1 << 32) + 15, (1 << 64) - 59]
    got = wwlib.cpu_residues(ps)
    for p, g in zip(ps, got):
        assert g == tuple(oracle.wieferich(p)) + tuple(oracle.wallsunsun(p)), p


def test_ladders_from_loosened_bounds():
    rng = random.Random(22)
    for p in sample_primes(rng, 300):
        want = wwlib.cpu_residues([p])
        lw, lf = lead(p >> 1), wss_lead(p)
        bounds = {(lw, lf), (63, 63), (max(lw, 31), max(lf, 31)), (max(lw, 32), max(lf, 32))}
        bounds.add((rng.randint(lw, 63), rng.randint(lf, 63)))
        for tw, tf in bounds:
            assert wwlib.cpu_residues([p], tw, tf) == want, (p, tw, tf)


def test_wss_table_start_lengths():
    """Five table steps are always allowed and six when the six leading bits of n below
    the bound give k <= 45 (V_k+1 below 2^64); both match the oracle, also for p below the
    table's entries (digits above p), and a refused six is an error."""
    rng = random.Random(24)
    ps = sample_primes(rng, 300) + [127, 131, 137, 1093, 3511, (1 << 64) - 59]
    sixes = fives_only = 0
    for p in ps:
        want = tuple(oracle.wieferich(p)) + tuple(oracle.wallsunsun(p))
        n = (p + 1 if p % 5 in (2, 3) else p - 1) // 2
        lf = lead(n)
        for tf in sorted({lf, min(lf + 1, 63), rng.randint(lf, 63), 63}):
            assert wwlib.cpu_residues([p], -1, tf, 5) == [want], (p, tf)
            if (n >> (tf - 5)) <= 45:
                sixes += 1
                assert wwlib.cpu_residues([p], -1, tf, 6) == [want], (p, tf)
            else:
                fives_only += 1
                with pytest.raises(wwlib.WWError):
                    wwlib.cpu_residues([p], -1, tf, 6)
    assert sixes > 200 and fives_only > 100
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_residues([1093], -1, -1, 4)


def test_windowed_prp_from_loosened_bounds():
    rng = random.Random(23)
    ns = [rng.getrandbits(rng.randint(8, 64)) | 1 for _ in range(3000)]
    ns += [(1 << t) + 1 for t in range(2, 64)]          # n - 1 a power of two
    ns += [3215031751, 4294967297, 2152302898747, 3474749660383, 25326001]
    for n in ns:
        if n < 3:
            continue
        want = oracle.strong_prp(2, n)
        e = n - 1
        lt, t = lead(e), (e & -e).bit_length() - 1
        assert wwlib.cpu_spsp2(n) == want, n
        top = rng.randint(lt, 63)
        tmax = rng.randint(t, top)
        assert wwlib.cpu_spsp2(n, top, tmax) == want, (n, top, tmax)
        assert wwlib.cpu_spsp2(n, 63, 63) == want, n


def test_bounds_below_the_leading_bit_are_refused():
    p = (1 << 63) + 25
    while not oracle.is_prp(p) or p % 5 == 0:
        p += 2
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_residues([p], 61, -1)
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_residues([p], -1, 64)
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_spsp2(p, 62, -1)
    with pytest.raises(wwlib.WWError):
        wwlib.cpu_residues([125], -1, -1)
