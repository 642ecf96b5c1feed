"""Unit checks of the device arithmetic's building blocks (csrc/ww_arith.h, compiled for
the host with g++): the 64x64 high-word products against 128-bit integer arithmetic, the
exact multiply-by-2^v with its quotient, and the digit-times-R^2 conversion into
Montgomery form mod p^2 against Python integers. Carry edge cases (all-ones halves,
zero halves, low words whose upper half is 0 or 2^32 - 1) are listed explicitly, since
random operands rarely reach them. No device needed."""
import os
import random
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "wwocl-boinc_amd", "csrc")

UNITS = r"""
#include <cstdio>
#include <cstdlib>
#include <random>
#include "ww_arith.h"
using namespace ww;
typedef unsigned __int128 u128x;

static const u64 kEdge[] = {0, 1, 2, 3, 0xffffffffull, 0x100000000ull, 0x1ffffffffull,
                            0xffffffff00000000ull, 0xfffffffe00000001ull, 0x8000000000000000ull,
                            0x7fffffffffffffffull, 0xffffffffffffffffull, 0x00000001ffffffffull};

static int check_mul_hi(u64 a, u64 b)
{
    const u128x P = (u128x)a * b;
    const u64 lo = (u64)P, hi = (u64)(P >> 64);
    if (mul_hi(a, b) != hi) return 1;
    if (mul_hi_known(a, b, lo) != hi) return 2;
    if (mul_hi_known(b, a, lo) != hi) return 3;
    return 0;
}

int main(int argc, char **argv)
{
    std::mt19937_64 g(argc > 1 ? strtoull(argv[1], 0, 10) : 1);
    for (u64 a : kEdge)
        for (u64 b : kEdge)
            if (int r = check_mul_hi(a, b)) { printf("mul_hi %d %llx %llx\n", r,
                                                      (unsigned long long)a, (unsigned long long)b); return 1; }
    for (int i = 0; i < 2000000; ++i) {
        u64 a = g(), b = g();
        switch (i & 3) { /* force halves to the carry extremes */
        case 1: a |= 0xffffffff00000000ull; b |= 0xffffffffull; break;
        case 2: a &= 0xffffffffull; break;
        case 3: b = 0 - b; break;
        }
        if (int r = check_mul_hi(a, b)) { printf("mul_hi %d %llx %llx\n", r,
                                                  (unsigned long long)a, (unsigned long long)b); return 1; }
    }
    /* y 2^v + c = quot n + r for y < n, v <= 31, c < 2^v */
    for (int i = 0; i < 1000000; ++i) {
        const int bits = 8 + (int)(g() % 57); /* moduli from 2^7 up: p >= 127 on the path */
        const u64 n = ((g() >> (64 - bits)) | (1ull << (bits - 1))) | 1;
        const u64 y = (i & 7) == 0 ? n - 1 : g() % n;
        const u32 v = (u32)(g() % 32);
        const u32 c = v ? (u32)(g() & ((1ull << v) - 1)) : 0;
        u32 quot;
        const u64 r = mulpow2_add_qr(y, v, c, n, 1.0 / (double)n, &quot);
        const u128x x = ((u128x)y << v) + c;
        if (r != (u64)(x % n) || quot != (u64)(x / n)) {
            printf("mulpow2 %llu %llu %u %u\n", (unsigned long long)y, (unsigned long long)n, v, c);
            return 1;
        }
    }
    /* Montgomery forms mod p^2 for the Python side: p, v, digits of m2_mul_digit(v, R^2) */
    for (int i = 0; i < 2000; ++i) {
        const int bits = 8 + (int)(g() % 57);
        const u64 p = ((g() >> (64 - bits)) | (1ull << (bits - 1))) | 1;
        if (p < 127 || p % 5 == 0) continue;
        const double pinv = 1.0 / (double)p;
        const m2 r2 = m2_r2(m2_one(p), p, pinv);
        const u64 v = (i & 3) == 0 ? ~0ull - (g() & 0xff) : g() >> (g() % 64);
        const m2 m = m2_mul_digit(v, r2, p, inv64(p));
        printf("%llu %llu %llu %llu %llu %llu\n", (unsigned long long)p, (unsigned long long)v,
               (unsigned long long)m.d0, (unsigned long long)m.d1, (unsigned long long)r2.d0,
               (unsigned long long)r2.d1);
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def units(tmp_path_factory):
    d = tmp_path_factory.mktemp("units")
    src = d / "units.cpp"
    src.write_text(UNITS)
    exe = d / "units"
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", str(exe), str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_products_quotients_and_montgomery_forms(units):
    r = subprocess.run([str(units), str(random.Random(3).getrandbits(32))], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stdout[-500:]
    lines = r.stdout.split("\n")
    rows = [tuple(int(x) for x in l.split()) for l in lines if l.strip()]
    assert len(rows) > 1000
    R = 1 << 64
    for p, v, d0, d1, r0, r1 in rows:
        p2 = p * p
        assert 0 <= d0 < p and 0 <= d1 < p and 0 <= r0 < p and 0 <= r1 < p, p
        assert r0 + p * r1 == R * R % p2, p              # R^2 mod p^2
        assert d0 + p * d1 == v * R % p2, (p, v)         # Montgomery form of v
