"""Writes tests/golden/golden.json: fixtures that pin the oracle.

Two independent sources, neither of them the oracle:
  * Python big-integer closed forms: s = 2^((p-1)/2) mod p^2 and F_{p-(p/5)} mod p^2
    split into base-p digits, plus a pow()-based strong probable-prime test;
  * the reference's own CPU arithmetic (prime.h, wwcpu.cpp) compiled into
    oracle/_ref/libwwref.so (oracle/Makefile, target `ref`): per-prime checksum slices
    and result flags, and range summaries of the loop of main.cpp:648-697 (including the
    full configurations [127, 10^8) and [2^50, 2^50 + 2^32)).

    python tests/golden/make_golden.py          (needs oracle/_ref built)
"""
from __future__ import annotations

import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402  (for the _ref binding only)

SMALL = [3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79,
         83, 89, 97, 101, 103, 107, 109, 113]


def strong_prp2(n: int) -> bool:
    """Miller-Rabin to base 2 straight from its definition."""
    d, t = n - 1, 0
    while d % 2 == 0:
        d //= 2
        t += 1
    x = pow(2, d, n)
    if x == 1 or x == n - 1:
        return True
    for _ in range(t - 1):
        x = x * x % n
        if x == n - 1:
            return True
    return False


def is_list_member(n: int) -> bool:
    return n % 2 == 1 and all(n % s for s in SMALL) and strong_prp2(n)


def fib_mod(n: int, m: int) -> int:
    """F_n mod m by fast doubling."""
    def pair(k):
        if k == 0:
            return 0, 1
        a, b = pair(k >> 1)
        c = a * ((2 * b - a) % m) % m
        d = (a * a + b * b) % m
        return (d, (c + d) % m) if k & 1 else (c, d)
    return pair(n)[0]


def closed_forms(p: int):
    p2 = p * p
    s = pow(2, (p - 1) // 2, p2)
    kro = -1 if p % 5 in (2, 3) else 1
    f = fib_mod(p - kro, p2)
    r = (1 << 64) % p2
    return {"p": p, "s0": s % p, "s1": s // p, "c20": f % p, "c21": f // p,
            "one0": r % p, "one1": r // p}


def members(lo: int, hi: int, k: int):
    out = []
    n = lo | 1
    while n < hi and len(out) < k:
        if is_list_member(n):
            out.append(n)
        n += 2
    return out


def main():
    if oracle.ref_lib() is None:
        sys.exit("oracle/_ref/libwwref.so is not built (make -C oracle ref)")
    rng = random.Random(20201226)
    ps = [1093, 3511, 495791, 42799, 49141, 88357, 90751, 104653, 1373653, 25326001,
          3215031751, 2276306935816523, 338772621946054253, 18446744073709551557]
    for lo in (127, (1 << 32) - 2000, 1 << 32, 1 << 50, (1 << 62) - 3000, 1 << 63,
               (1 << 64) - 3000):
        ps += members(lo, min(lo + 3000, (1 << 64) - 1), 12)
    for bits in range(8, 65):
        for _ in range(4):
            lo = max(127, rng.getrandbits(bits) | (1 << (bits - 1)))
            ps += members(lo, min(lo + 3000, (1 << 64) - 1), 2)
    ps = sorted(set(ps))
    primes = []
    for p in ps:
        row = closed_forms(p)
        fw, sw, fs, ss = oracle.ref_slices(p)
        row.update({"ref_wflag": fw, "ref_wslice": sw, "ref_sflag": fs, "ref_sslice": ss})
        primes.append(row)

    ranges = []
    spans = [(127, 1_000_000), (1_000_000, 2_000_000), (10**9, 10**9 + 10**6),
             ((1 << 32) - (1 << 19), (1 << 32) + (1 << 19)), (1 << 50, (1 << 50) + (1 << 20)),
             (1 << 63, (1 << 63) + (1 << 20)), ((1 << 64) - (1 << 20), (1 << 64) - 1),
             (2276306935816523 - 5000, 2276306935816523 + 5000),
             (338772621946054253, 338772621946054253 + 5000)]
    for lo, hi in spans:
        prp, cks, flags = oracle.ref_range(lo, hi, 8)
        ranges.append({"lo": lo, "hi": hi, "prp_count": prp, "checksum": cks,
                       "cpu_result_count": flags})

    # the full parity configurations: [127, 10^8) and [2^50, 2^50 + 2^32)
    full = []
    for name, lo, hi in (("127_to_1e8", 127, 10**8), ("2p50_w2p32", 1 << 50, (1 << 50) + (1 << 32))):
        prp, cks, flags = oracle.ref_range(lo, hi, os.cpu_count() or 1)
        full.append({"name": name, "lo": lo, "hi": hi, "prp_count": prp, "checksum": cks,
                     "cpu_result_count": flags})

    spsp_composites = [n for n in range(3, 200_001, 2)
                       if strong_prp2(n) and any(n % q == 0 for q in range(3, int(n**0.5) + 1, 2))]
    fixture = {
        "about": "closed forms (Python big ints) and reference CPU path (oracle/_ref)",
        "primes": primes,
        "ranges": ranges,
        "full_ranges": full,
        "spsp2_composites_below_200001": spsp_composites,
        "pi": {"1e6": 78498, "1e7": 664579, "1e8": 5761455},
    }
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(fixture, f, indent=0)
    print(f"wrote {len(primes)} primes, {len(ranges)} ranges, {len(full)} full configurations")


if __name__ == "__main__":
    main()
