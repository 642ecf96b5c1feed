/*
 * ww_selftest.cpp -- the host-side checks of the drop-in in a formulation of their own,
 * independent of the device arithmetic (ww_arith.h): plain 128-bit products reduced by
 * division, no Montgomery form, no windows, no Lucas V ladder. The startup self-test
 * compares the device against this code, as the reference compares its kernels against
 * the separate CPU functions wieferichCPU / wallsunsunCPU (main.cpp:648-697), so a fault
 * in the device formulation cannot hide behind the same fault on the host.
 *   ww_cpu_is_prime   isPrime, main.cpp:301-324 (Miller-Rabin, bases 2..23 or 2..37)
 *   ww_cpu_is_prp     isPRP, main.cpp:245-291 (no factor 3..113, base-2 strong PRP)
 *   ww_cpu_range      testKernels' CPU loop, main.cpp:648-697
 */
#include <stdint.h>

#include "../../include/ww_api.h"

namespace {

typedef unsigned __int128 u128;
typedef uint64_t u64;

u64 mulmod(u64 a, u64 b, u64 n) { return (u64)((u128)a * b % n); }

u64 addmod(u64 a, u64 b, u64 n) /* a, b < n */
{
    return a >= n - b ? a - (n - b) : a + b;
}

u64 submod(u64 a, u64 b, u64 n) /* a, b < n */
{
    return a >= b ? a - b : a + (n - b);
}

u64 powmod(u64 b, u64 e, u64 n)
{
    u64 r = 1 % n, x = b % n;
    for (; e; e >>= 1) {
        if (e & 1) r = mulmod(r, x, n);
        x = mulmod(x, x, n);
    }
    return r;
}

/* Miller-Rabin to base a of odd n > 2, by the definition: n - 1 = d 2^t, a^d = 1 or
   a^(d 2^s) = -1 for some s < t. */
bool strong_prp(u64 a, u64 n)
{
    u64 d = n - 1;
    int t = 0;
    while ((d & 1) == 0) {
        d >>= 1;
        ++t;
    }
    u64 x = powmod(a, d, n);
    if (x == 1 || x == n - 1) return true;
    for (int s = 1; s < t; ++s) {
        x = mulmod(x, x, n);
        if (x == n - 1) return true;
    }
    return false;
}

bool no_factor_to_113(u64 n)
{
    static const unsigned small[29] = {3,  5,  7,  11, 13, 17, 19, 23, 29, 31,
                                       37, 41, 43, 47, 53, 59, 61, 67, 71, 73,
                                       79, 83, 89, 97, 101, 103, 107, 109, 113};
    for (unsigned d : small)
        if (n % d == 0) return false;
    return true;
}

/* A residue mod p^2 as two base-p digits: v = lo + p hi, 0 <= lo, hi < p. */
struct P2 {
    u64 lo, hi;
};

/* x y mod p^2 = x0 y0 + p (x0 y1 + x1 y0) mod p^2, with x0 y0 = a + p b (b < p). */
P2 mul2(P2 x, P2 y, u64 p)
{
    const u128 t = (u128)x.lo * y.lo;
    P2 r;
    r.lo = (u64)(t % p);
    const u64 b = (u64)(t / p);
    r.hi = addmod(addmod(b, mulmod(x.lo, y.hi, p), p), mulmod(x.hi, y.lo, p), p);
    return r;
}

P2 add2(P2 x, P2 y, u64 p)
{
    P2 r;
    const bool carry = x.lo >= p - y.lo;
    r.lo = carry ? x.lo - (p - y.lo) : x.lo + y.lo;
    r.hi = addmod(addmod(x.hi, y.hi, p), carry ? 1 : 0, p);
    return r;
}

P2 sub2(P2 x, P2 y, u64 p)
{
    P2 r;
    const bool borrow = x.lo < y.lo;
    r.lo = borrow ? x.lo + (p - y.lo) : x.lo - y.lo;
    r.hi = submod(submod(x.hi, y.hi, p), borrow ? 1 : 0, p);
    return r;
}

/* 2^((p-1)/2) mod p^2, left to right over the bits of the exponent. */
P2 wieferich(u64 p)
{
    const u64 e = p >> 1;
    P2 x = {1 % p, 0};
    for (int i = 63; i >= 0; --i) {
        x = mul2(x, x, p);
        if ((e >> i) & 1) x = add2(x, x, p);
    }
    return x;
}

/* F_m mod p^2 for m = p - (p/5), by the Fibonacci pair (F_k, F_k+1):
   F_2k = F_k (2 F_k+1 - F_k), F_2k+1 = F_k^2 + F_k+1^2. */
P2 wallsunsun(u64 p)
{
    const unsigned r5 = (unsigned)(p % 5);
    const u64 m = (r5 == 2 || r5 == 3) ? p + 1 : p - 1; /* p < 2^64 - 1 odd, so no wrap */
    P2 f = {0, 0}, g = {1 % p, 0};                       /* F_0, F_1 */
    for (int i = 63; i >= 0; --i) {
        const P2 two_g = add2(g, g, p);
        const P2 f2 = mul2(f, sub2(two_g, f, p), p); /* F_2k */
        const P2 g2 = add2(mul2(f, f, p), mul2(g, g, p), p); /* F_2k+1 */
        if ((m >> i) & 1) {
            f = g2;
            g = add2(f2, g2, p); /* F_2k+2 */
        } else {
            f = f2;
            g = g2;
        }
    }
    return f;
}

} // namespace

extern "C" {

int ww_cpu_is_prime(uint64_t n)
{
    static const u64 bases[12] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    if ((n & 1) == 0) return 0;
    int nb = n < 3825123056546413051ull ? 9 : 12;
    for (int i = 0; i < nb; ++i)
        if (!strong_prp(bases[i], n)) return 0;
    return 1;
}

int ww_cpu_is_prp(uint64_t n)
{
    if ((n & 1) == 0) return 0;
    return no_factor_to_113(n) && strong_prp(2, n);
}

int ww_cpu_range(uint64_t low, uint64_t high, uint64_t out[3])
{
    if (!out) return WW_E_ARG;
    const u64 T = 1000; /* SPECIAL_THRESHOLD, WW.h:9 */
    u64 prp = 0, sum = 0, flags = 0;
    u64 n = low | 1;
    if (n >= low) {
        for (; n < high; n += 2) {
            if (n > 113 && no_factor_to_113(n) && strong_prp(2, n)) {
                ++prp;
                const P2 s = wieferich(n);
                const P2 f = wallsunsun(n);
                sum += (s.hi & 0xffffffffu) + (f.hi & 0xffffffffu);
                /* wieferichCPU / wallsunsunCPU return values (wwcpu.cpp:290-303, 382-395) */
                if ((s.hi == 0 && s.lo == 1) || (s.hi == n - 1 && s.lo == n - 1) || s.hi <= T ||
                    s.hi >= n - T)
                    ++flags;
                if (f.hi <= T || f.hi >= n - T) ++flags;
            }
            if (n + 2 < n) break;
        }
    }
    out[0] = prp;
    out[1] = sum;
    out[2] = flags;
    return WW_OK;
}

} // extern "C"
