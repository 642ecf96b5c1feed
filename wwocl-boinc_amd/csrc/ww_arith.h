/*
 * ww_arith.h -- exact integer arithmetic of the search path, written for CDNA4 VGPRs.
 *
 * Compiles as device code (hipcc) and as host code (g++ or hipcc host pass), so the
 * very same routines run the host-side self-test and result validation that the
 * reference performs on the CPU (main.cpp:590-738, 768-956).
 *
 * Layout: 64-bit values are pairs of 32-bit VGPRs. Every 32x32->64 partial product is
 * written as (u64)a*b + c with 32-bit a, b so the backend emits v_mad_u64_u32
 * (one MUL32 each); 64-bit carries become v_add_co_u32 / v_addc_co_u32 pairs.
 *
 * Residues mod p^2 are two base-p digits (d0 + p*d1, 0 <= di < p) in Montgomery form
 * with factor 2^64 -- the representation of the reference (wwcpu.cpp:104-250;
 * Dorais & Klyve, J. Integer Seq. 14 (2011)). All outputs are exact, so any exact
 * evaluation order yields the reference's values; the Wall-Sun-Sun ladder here uses
 * the Lucas sequence V_n(3, 1) (one square + one product per exponent bit) instead of the
 * reference's Fibonacci pair (two squares + one product, wallsunsun.cl:274-294).
 */
#ifndef WW_ARITH_H
#define WW_ARITH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define WW_HD __host__ __device__ __forceinline__
#else
#define WW_HD static inline
#endif

/* WW_LOOP_LOCAL(x): an empty asm that makes x opaque. At the top of a loop body: loop-invariant
   64-bit operands otherwise have their 32-bit halves hoisted as zero-extended 64-bit values,
   and instruction selection, which sees one block at a time, then multiplies by the known-
   zero high halves too (8 v_mad_u64_u32 for a 4-MUL32 high product). On a value like 2^64 - p:
   keeps "x + (2^64 - p)" (one 64-bit add) from being rewritten as "x - p" (two). */
#if defined(__HIP_DEVICE_COMPILE__)
#define WW_LOOP_LOCAL(x) asm volatile("" : "+v"(x))
#else
#define WW_LOOP_LOCAL(x) ((void)0)
#endif

namespace ww {

typedef uint64_t u64;
typedef uint32_t u32;

struct u128 {
    u64 lo, hi;
};

/* Residue mod p^2: d0 + p*d1, both digits in [0, p). */
struct m2 {
    u64 d0, d1;
};

WW_HD u64 mad32(u32 a, u32 b, u64 c) { return (u64)a * b + c; }

WW_HD int clz64(u64 x) { return __builtin_clzll(x); }
WW_HD int ctz64(u64 x) { return __builtin_ctzll(x); }

/* 64x64 -> 128, four MUL32. No partial sum can overflow 64 bits:
   (2^32-1)^2 + 2*(2^32-1) = 2^64-1. */
WW_HD u128 mul_wide(u64 a, u64 b)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 t0 = (u64)a0 * b0;
    u64 t1 = mad32(a0, b1, t0 >> 32);
    u64 t2 = mad32(a1, b0, (u32)t1);
    u64 t3 = mad32(a1, b1, (t1 >> 32) + (t2 >> 32));
    u128 r;
    r.lo = (u64)(u32)t0 | (t2 << 32);
    r.hi = t3;
    return r;
}

/* Low 64 bits of a*b. On the device as three v_mad_u64_u32: a0*b0, then the two cross
   products accumulated into its high word. Only the low words of the cross MADs matter,
   and they are emitted as written (the backend would otherwise narrow them to two
   v_mul_lo_u32 and a v_add3_u32, four VOP3 operations where three MADs do). */
WW_HD u64 mul_lo64(u64 a, u64 b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 t = (u64)a0 * b0;
    u64 s, r, c1, c2;
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(s), "=s"(c1) : "v"(a0), "v"(b1), "v"(t >> 32));
    asm("v_mad_u64_u32 %0, %1, %2, %3, %4" : "=v"(r), "=s"(c2) : "v"(a1), "v"(b0), "v"(s));
    (void)c1;
    (void)c2;
    return ((u64)(u32)r << 32) | (u32)t;
#else
    return a * b;
#endif
}

/* High 64 bits of a*b. */
WW_HD u64 mul_hi(u64 a, u64 b)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 t0 = (u64)a0 * b0;
    u64 t1 = mad32(a0, b1, t0 >> 32);
    u64 t2 = mad32(a1, b0, (u32)t1);
    return mad32(a1, b1, (t1 >> 32) + (t2 >> 32));
}

/* High word of a*b when its low word lo is already known, as in every Montgomery
   reduction here: u = t_lo * p^-1 makes the low word of u*p equal t_lo. With
   a*b = a1 b1 2^64 + (a1 b0 + a0 b1) 2^32 + a0 b0, the middle column plus the high half h
   of a0 b0 has its low 32 bits fixed to hi32(lo), so its carry into the high word is
   k = ceil((a1 b0 + a0 b1 - hi32(lo)) / 2^32) = (a1 b0 + a0 b1 + ~hi32(lo)) >> 32 for every
   h in [0, 2^32). Three 32-bit products instead of four, and no partial sum overflows:
   a1 b0 + ~hi32(lo) <= 2^64 - 2^32, a0 b1 + 2^32 - 1 < 2^64, a1 b1 + 2 (2^32 - 1) < 2^64. */
WW_HD u64 mul_hi_known(u64 a, u64 b, u64 lo)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 s = mad32(a1, b0, (u64)(~(u32)(lo >> 32)));
    u64 t = mad32(a0, b1, (u32)s);
    return mad32(a1, b1, s >> 32) + (t >> 32);
}

/* a^2 -> 128: the cross product a0*a1 is formed by two accumulating MADs, which
   doubles it without a 65-bit intermediate. */
WW_HD u128 sqr_wide(u64 a)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    u64 t0 = (u64)a0 * a0;
    u64 t1 = mad32(a0, a1, t0 >> 32);     /* a0a1 + hi(a0^2)        */
    u64 t2 = mad32(a0, a1, (u32)t1);      /* a0a1 + lo(t1)          */
    u64 t3 = mad32(a1, a1, (t1 >> 32) + (t2 >> 32));
    u128 r;
    r.lo = (u64)(u32)t0 | (t2 << 32);
    r.hi = t3;
    return r;
}

/* a*b + c + d, exact when the result is below 2^128. The addends enter the schoolbook
   columns as 32-bit halves, two per MAD at most, so no partial sum exceeds
   (2^32-1)^2 + 2(2^32-1) = 2^64-1 and no separate carry chain is needed. */
WW_HD u128 mul_wide_add2(u64 a, u64 b, u64 c, u64 d)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 t0 = mad32(a0, b0, (u64)(u32)c + (u32)d);
    u64 t1 = mad32(a0, b1, (t0 >> 32) + (c >> 32));
    u64 t2 = mad32(a1, b0, (u64)(u32)t1 + (d >> 32));
    u64 t3 = mad32(a1, b1, (t1 >> 32) + (t2 >> 32));
    u128 r;
    r.lo = (u64)(u32)t0 | (t2 << 32);
    r.hi = t3;
    return r;
}

/* a*b + c (always below 2^128). */
WW_HD u128 mul_wide_add(u64 a, u64 b, u64 c)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 t0 = mad32(a0, b0, (u32)c);
    u64 t1 = mad32(a0, b1, (t0 >> 32) + (c >> 32));
    u64 t2 = mad32(a1, b0, (u32)t1);
    u64 t3 = mad32(a1, b1, (t1 >> 32) + (t2 >> 32));
    u128 r;
    r.lo = (u64)(u32)t0 | (t2 << 32);
    r.hi = t3;
    return r;
}

/* a^2 + c (always below 2^128). */
WW_HD u128 sqr_wide_add(u64 a, u64 c)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    u64 t0 = mad32(a0, a0, (u32)c);
    u64 t1 = mad32(a0, a1, (t0 >> 32) + (c >> 32));
    u64 t2 = mad32(a0, a1, (u32)t1);
    u64 t3 = mad32(a1, a1, (t1 >> 32) + (t2 >> 32));
    u128 r;
    r.lo = (u64)(u32)t0 | (t2 << 32);
    r.hi = t3;
    return r;
}

/* ---------------------------------------------------------------------------
 * mod n (64-bit) helpers. Borrows and carries come from the overflow builtins so the
 * backend reuses the carry-out of v_sub_co/v_add_co instead of comparing again.
 * ------------------------------------------------------------------------- */

/* x - y, and whether it borrowed. Written as two 32-bit subtractions: the backend then
   takes the borrow from v_subb_co's carry-out, where the 64-bit builtin gets a second
   64-bit compare (the high word borrows when x_hi < y_hi, or when they are equal and the
   low word borrowed). */
WW_HD u64 sub_b(u64 x, u64 y, bool *borrow)
{
    u32 rl, t, rh;
    bool b1 = __builtin_sub_overflow((u32)x, (u32)y, &rl);
    bool b2 = __builtin_sub_overflow((u32)(x >> 32), (u32)(y >> 32), &t);
    bool b3 = __builtin_sub_overflow(t, (u32)b1, &rh);
    *borrow = b2 | b3;
    return ((u64)rh << 32) | rl;
}

/* a + b + c for a carry c in {0, 1}, and the carry out (one v_addc_co on the device). */
WW_HD u32 addc32(u32 a, u32 b, bool *c)
{
    u32 t, r;
    bool c1 = __builtin_add_overflow(a, b, &t);
    bool c2 = __builtin_add_overflow(t, (u32)*c, &r);
    *c = c1 | c2;
    return r;
}

/* a + e as four 32-bit words, with the carry out of bit 127: a v_add_co / v_addc_co chain
   whose last carry is the result, instead of 64-bit adds each followed by a compare. */
WW_HD u128 add128_c(u128 a, u128 e, bool *carry)
{
    bool c = false;
    u32 r0 = addc32((u32)a.lo, (u32)e.lo, &c);
    u32 r1 = addc32((u32)(a.lo >> 32), (u32)(e.lo >> 32), &c);
    u32 r2 = addc32((u32)a.hi, (u32)e.hi, &c);
    u32 r3 = addc32((u32)(a.hi >> 32), (u32)(e.hi >> 32), &c);
    *carry = c;
    u128 r;
    r.lo = ((u64)r1 << 32) | r0;
    r.hi = ((u64)r3 << 32) | r2;
    return r;
}

/* x + y mod p for x, y < p (prime.h:34-43). */
WW_HD u64 add_mod(u64 x, u64 y, u64 p)
{
    u64 c = (x >= p - y) ? p : 0;
    return x + y - c;
}

/* 2x mod p for x < p, with np = 2^64 - p: 2x >= p exactly when the shift carries out or
   the shifted value is >= p, and adding np subtracts p modulo 2^64 (one 64-bit add). */
WW_HD u64 dbl_mod(u64 x, u64 p, u64 np)
{
    u64 d = x << 1;
    bool ge = (x >> 63) | (d >= p);
    return d + (ge ? np : 0);
}

/* x - y mod p for x < p, y <= p. */
WW_HD u64 sub_mod(u64 x, u64 y, u64 p)
{
    bool b;
    u64 r = sub_b(x, y, &b);
    return b ? r + p : r;
}

/* a - b - c for a borrow c in {0, 1}, and the borrow out (one v_subb_co on the device). */
WW_HD u32 subb32(u32 a, u32 b, bool *c)
{
    u32 t, r;
    bool b1 = __builtin_sub_overflow(a, b, &t);
    bool b2 = __builtin_sub_overflow(t, (u32)*c, &r);
    *c = b1 | b2;
    return r;
}

/* x - (y + c) mod p for x < p, y < p and a borrow c in {0, 1}: the borrow enters the
   subtraction's carry chain instead of being added to y first. */
WW_HD u64 sub_mod_c(u64 x, u64 y, bool c, u64 p)
{
    u32 r0 = subb32((u32)x, (u32)y, &c);
    u32 r1 = subb32((u32)(x >> 32), (u32)(y >> 32), &c);
    u64 r = ((u64)r1 << 32) | r0;
    return c ? r + p : r;
}

/* p^-1 mod 2^64 for odd p: Newton from the 5-bit seed (3p) xor 2, three 32-bit steps
   then one 64-bit step (same value as invert(), prime.h:7-12). */
WW_HD u64 inv64(u64 p)
{
    u32 p32 = (u32)p;
    u32 x = (3u * p32) ^ 2u;
    x *= 2u - p32 * x;
    x *= 2u - p32 * x;
    x *= 2u - p32 * x;
    u64 X = x;
    X *= 2u - p * X;
    return X;
}

/* floor(2^64 / n) and 2^64 mod n for odd n >= 2^32, without a 64-bit division: a
   double estimate of the quotient is within one of the truth and is corrected by the
   exact signed remainder 2^64 - Qe*n. */
WW_HD void pow64_divmod(u64 n, u64 *quot, u64 *rem)
{
    if (n >> 63) { /* 2^63 < n < 2^64 */
        *quot = 1;
        *rem = 0 - n;
        return;
    }
    double qd = 18446744073709551616.0 / (double)n;
    u64 qe = (u64)qd; /* qd < 2^32 + 1 */
    u128 pr = mul_wide(qe, n);
    u64 lo = 0 - pr.lo;
    u64 hi = 1 - pr.hi - (pr.lo != 0 ? 1 : 0); /* (2^64 - qe*n) >> 64, two's complement */
    if ((int64_t)hi < 0) {                      /* qe one too large */
        *quot = qe - 1;
        *rem = lo + n;
    } else if (lo >= n) {                       /* qe one too small */
        *quot = qe + 1;
        *rem = lo - n;
    } else {
        *quot = qe;
        *rem = lo;
    }
}

/* 2^64 mod n (Montgomery 1, montInit prime.h:49 / getsegprimes.cl:107). */
WW_HD u64 mont_one(u64 n)
{
    if (n >> 32) {
        u64 qq, r;
        pow64_divmod(n, &qq, &r);
        return r;
    }
    return (0 - n) % n;
}

/* Montgomery square a^2/2^64 mod n with q = n^-1 mod 2^64 (prime.h:15-31). */
WW_HD u64 mont_sqr(u64 a, u64 n, u64 q)
{
    u128 t = sqr_wide(a);
    u64 m = mul_lo64(t.lo, q);
    u64 mp = mul_hi_known(m, n, t.lo);
    bool b;
    u64 r = sub_b(t.hi, mp, &b);
    return b ? r + n : r;
}
/* Montgomery product a*b/2^64 mod n (prime.h:15-31). */
WW_HD u64 mont_mul(u64 a, u64 b, u64 n, u64 q)
{
    u128 t = mul_wide(a, b);
    u64 m = mul_lo64(t.lo, q);
    u64 mp = mul_hi_known(m, n, t.lo);
    u64 r = t.hi - mp;
    return t.hi < mp ? r + n : r;
}

/* A 64-bit integer as the nearest double: both 32-bit halves convert exactly and one FMA
   rounds their sum once. */
WW_HD double u64_to_double(u64 y)
{
    return fma((double)(u32)(y >> 32), 4294967296.0, (double)(u32)y);
}

/* y * 2^v + c = quot * n + r for y < n, v <= 31 and c < 2^v; returns r and the quotient
   (below 2^v). With x = y 2^v + c < n 2^31, x and 1/n in double precision carry a
   relative error below 2^-50.6 together (three roundings of x, two of 1/n, the FMA's
   own), an absolute error below 2^-19.6 at quotients below 2^31, so
   floor(x * (1/n) + 2^-18) lies in (x/n, x/n + 1): it is the quotient or one more. The
   remainder x - qe n then lies in [-n, n), and when the 96-bit subtraction borrows, one
   addition of n and a decrement of the quotient finish it. */
WW_HD u64 mulpow2_add_qr(u64 y, u32 v, u32 c, u64 n, double ninv, u32 *quot)
{
    u32 qe = (u32)fma(ldexp(u64_to_double(y), (int)v) + (double)c, ninv, 0x1p-18);
    /* x = hi:lo */
    u64 sh = y << v;
    bool cy = false;
    u32 lo0 = addc32((u32)sh, c, &cy);
    u32 lo1 = addc32((u32)(sh >> 32), 0u, &cy);
    u32 hi = addc32(((u32)(y >> 32) >> 1) >> (31 - v), 0u, &cy); /* y >> (64 - v) */
    /* qe n = m1:m0 (96 bits) */
    u64 m0 = (u64)qe * (u32)n;
    u64 m1 = mad32(qe, (u32)(n >> 32), m0 >> 32);
    bool b = false;
    u32 r0 = subb32(lo0, (u32)m0, &b);
    u32 r1 = subb32(lo1, (u32)m1, &b);
    (void)subb32(hi, (u32)(m1 >> 32), &b); /* b: x < qe n */
    u64 r = ((u64)r1 << 32) | r0;
    *quot = qe - (u32)b;
    return b ? r + n : r;
}

/* y * 2^v mod n for y < n and v <= 31. */
WW_HD u64 mulpow2_mod(u64 y, u32 v, u64 n, double ninv)
{
    u32 quot;
    return mulpow2_add_qr(y, v, 0, n, ninv, &quot);
}

/* Base-2 strong probable-prime test of odd n > 1 (getsegprimes.cl:118-164,
   prime.h:66-112), q = n^-1 mod 2^64, one = 2^64 mod n.
   With n - 1 = d 2^t the test asks for x_0 = 2^d = 1 or x_s = x_0^(2^s) = -1 for some
   s < t. The bits of n - 1 from `top` down to `tmax` yield 2^((n-1) >> tmax): the first
   five (or all, if fewer) as one multiplication of 1 by 2^v, then the rest in windows of
   squarings followed by one multiplication by 2^v, a short window of (len - 5) mod 5 bits
   first if that is not 0, then windows of 5. The bits below tmax follow one by one, and
   x_0 .. x_{t-1} appear there (or x_0 at the end of the windows when t = tmax). Any
   top >= the leading bit of n - 1 and any tmax in [t, top] give the same answer, so a
   wave can pass common (uniform) values and run one loop for all lanes. */
WW_HD bool spsp2_window(u64 n, u64 q, u64 one, int top, int tmax)
{
    const int W = 5;
    u64 e = n - 1;
    int t = ctz64(e);
    double ninv = 1.0 / (double)n;
    u64 mone = n - one;
    int len = top - tmax + 1;
    int first = len < W ? len : W;
    int pos = top - first + 1;
    u64 a = mulpow2_mod(one, (u32)(e >> pos) & ((1u << first) - 1u), n, ninv);
    while (pos > tmax) {
        const int w = (pos - tmax) % W ? (pos - tmax) % W : W; /* uniform: short one first */
        pos -= w;
        for (int k = 0; k < w; ++k) a = mont_sqr(a, n, q);
        a = mulpow2_mod(a, (u32)(e >> pos) & ((1u << w) - 1u), n, ninv);
    }
    bool ok = (t == tmax) & ((a == one) | (a == mone));
    for (int i = tmax - 1; i >= 1; --i) {
        a = mont_sqr(a, n, q);
        if ((e >> i) & 1) a = add_mod(a, a, n);
        ok |= ((i == t) & ((a == one) | (a == mone))) | ((i < t) & (a == mone));
    }
    return ok;
}

/* The same with the exact bounds of this n. */
WW_HD bool spsp2(u64 n, u64 q, u64 one)
{
    u64 e = n - 1;
    return spsp2_window(n, q, one, 63 - clz64(e), ctz64(e));
}

/* The deep sieve's start positions in a segment of odd numbers G + 2i (G odd) for one
   prime ell < 2^16 with m = ell * lcm(2, ord_ell(2)) < 2^32: i_mark in [0, ell) is the
   first odd multiple of ell, i_resc in [0, m/2) the first member of the pseudoprime
   class ell mod m. Integer only: G = hi 2^32 + lo is congruent mod m to
   hi (2^32 mod m) + lo < 2^64 (one MAD), reduced by a 64-bit Barrett step with
   bar_m = floor(2^64 / m); then 32-bit Barrett steps with bar_ell = floor(2^32 / ell)
   give G mod ell and i_mark = -G (ell + 1)/2 mod ell. Each Barrett quotient is the true
   one or one less, so one conditional subtraction finishes each step. */
WW_HD void deep_offsets(u64 G, u32 ell, u32 m, u32 c32m, u32 bar_ell, u64 bar_m,
                        u32 *i_mark, u32 *i_resc)
{
    u64 B = mad32((u32)(G >> 32), c32m, (u32)G); /* = G mod m, below 2^64 */
    u64 rm = B - mul_hi(B, bar_m) * m;
    if (rm >= m) rm -= m;
    u32 r = (u32)rm;
    u32 rl = r - (u32)(((u64)r * bar_ell) >> 32) * ell;
    if (rl >= ell) rl -= ell;
    u32 x = (rl ? ell - rl : 0u) * ((ell + 1) >> 1); /* < 2^31 */
    u32 im = x - (u32)(((u64)x * bar_ell) >> 32) * ell;
    if (im >= ell) im -= ell;
    *i_mark = im;
    *i_resc = (ell >= r ? ell - r : ell + m - r) >> 1; /* (ell - G) mod m is even */
}

/* ---------------------------------------------------------------------------
 * mod p^2 residues (two base-p digits, Montgomery factor 2^64)
 * ------------------------------------------------------------------------- */

/* 2^64 mod p^2 as digits (wwcpu.cpp:109-125). */
WW_HD m2 m2_one(u64 p)
{
    m2 r;
    if (p >> 32) {
        pow64_divmod(p, &r.d1, &r.d0); /* 2^64 = d0 + p*d1 with d1 = floor(2^64/p) */
    } else {
        u64 p2 = p * p;
        u64 rr = (0 - p2) % p2;
        r.d0 = rr % p;
        r.d1 = rr / p;
    }
    return r;
}

/* x + y mod p^2 (wwcpu.cpp:241-250). */
WW_HD m2 m2_add(m2 x, m2 y, u64 p)
{
    bool c = x.d0 >= p - y.d0;
    m2 r;
    r.d0 = x.d0 + y.d0 - (c ? p : 0);
    r.d1 = add_mod(x.d1 + (c ? 1 : 0), y.d1, p);
    return r;
}

/* x - y mod p^2. */
WW_HD m2 m2_sub(m2 x, m2 y, u64 p)
{
    bool b = x.d0 < y.d0;
    m2 r;
    r.d0 = x.d0 - y.d0 + (b ? p : 0);
    r.d1 = sub_mod_c(x.d1, y.d1, b, p);
    return r;
}

/* -x mod p^2. */
WW_HD m2 m2_neg(m2 x, u64 p)
{
    m2 z;
    z.d0 = 0;
    z.d1 = 0;
    return m2_sub(z, x, p);
}

/* Montgomery reduction tail shared by square and product: t = x0*y0 (128 bits) and
   the cross term c (129 bits: c_lo, c_hi, c_top) become the digits of
   (x*y)/2^64 mod p^2 (wwcpu.cpp:143-166, 213-236). */
WW_HD m2 m2_redc(u128 t, u64 c_lo, u64 c_hi, u64 c_top, u64 p, u64 q)
{
    u64 u0 = mul_lo64(t.lo, q);
    u64 v1 = mul_hi_known(p, u0, t.lo);
    /* tp = c + u0 */
    u64 tlo = c_lo + u0;
    u64 cy = tlo < u0 ? 1 : 0;
    u64 thi = c_hi + cy;
    u64 top = c_top | ((thi < cy) ? 1 : 0);
    u64 t1p = (top | (thi >= p ? 1 : 0)) ? thi - p : thi; /* tp >> 64 < 2p */
    u64 up0 = mul_lo64(tlo, q);
    u64 v1p = mul_hi_known(p, up0, tlo);
    bool b = t.hi < v1;
    m2 r;
    r.d0 = t.hi - v1 + (b ? p : 0);
    r.d1 = sub_mod_c(t1p, v1p, b, p);
    return r;
}

/* x^2 / 2^64 mod p^2 (wwcpu.cpp:140-167). The cross term 2 x0 x1 enters only through
   its REDC into the high digit, so only its value mod p matters: it is formed as
   x0 * (2 x1 mod p) + u0 <= (p-1)^2 + 2^64 - 1 < 2^128, whose high word is already below
   p. That drops the 129-bit sum and its conditional subtraction of the general product. */
WW_HD m2 m2_sqr(m2 x, u64 p, u64 q, u64 np)
{
    u128 t = sqr_wide(x.d0);
    u64 u0 = mul_lo64(t.lo, q);
    u64 v1 = mul_hi_known(p, u0, t.lo);
    u128 c = mul_wide_add(x.d0, dbl_mod(x.d1, p, np), u0); /* high word < p */
    u64 up0 = mul_lo64(c.lo, q);
    u64 v1p = mul_hi_known(p, up0, c.lo);
    bool b;
    u64 z0 = sub_b(t.hi, v1, &b);
    m2 r;
    r.d0 = b ? z0 + p : z0;
    r.d1 = sub_mod_c(c.hi, v1p, b, p);
    return r;
}

/* x*y / 2^64 mod p^2 (wwcpu.cpp:210-237). */
WW_HD m2 m2_mul(m2 x, m2 y, u64 p, u64 q)
{
    u128 t = mul_wide(x.d0, y.d0);
    u128 a = mul_wide(x.d0, y.d1);
    u128 b = mul_wide(x.d1, y.d0);
    u64 lo = a.lo + b.lo;
    u64 c1 = lo < a.lo ? 1 : 0;
    u64 hi = a.hi + b.hi;
    u64 top = hi < a.hi ? 1 : 0;
    hi += c1;
    top |= (hi < c1) ? 1 : 0;
    return m2_redc(t, lo, hi, top, p, q);
}

/* REDC(x^2 + k) for a constant residue k = k0 + p*k1 (digits in [0, p)): with
   k = -c*R^2 mod p^2 this is the Montgomery square minus c, so the ladder's "- 2" costs
   two 32-bit additions instead of a mod p^2 subtraction. Bounds: x0^2 + k0 < p^2 and
   x0*(2x1 mod p) + u0 + k1 < p*2^64, so both high words are below p as in m2_sqr. */
WW_HD m2 m2_sqr_add(m2 x, m2 k, u64 p, u64 q, u64 np)
{
    u128 t = sqr_wide_add(x.d0, k.d0);
    u64 u0 = mul_lo64(t.lo, q);
    u64 v1 = mul_hi_known(p, u0, t.lo);
    u128 c = mul_wide_add2(x.d0, dbl_mod(x.d1, p, np), u0, k.d1);
    u64 up0 = mul_lo64(c.lo, q);
    u64 v1p = mul_hi_known(p, up0, c.lo);
    bool b;
    u64 z0 = sub_b(t.hi, v1, &b);
    m2 r;
    r.d0 = b ? z0 + p : z0;
    r.d1 = sub_mod_c(c.hi, v1p, b, p);
    return r;
}

/* v*y / 2^64 mod p^2 for any 64-bit v (the product with a zero high digit; v may exceed
   p). v*y0 and the cross term v*y1 + u0 are both at most (2^64 - 1) p < p*2^64, so their
   high words are below p as in m2_sqr. With y = R^2 mod p^2 (R = 2^64) this is v's
   Montgomery form, v*R mod p^2. */
WW_HD m2 m2_mul_digit(u64 v, m2 y, u64 p, u64 q)
{
    u128 t = mul_wide(v, y.d0);
    u64 u0 = mul_lo64(t.lo, q);
    u64 v1 = mul_hi_known(p, u0, t.lo);
    u128 c = mul_wide_add(v, y.d1, u0);
    u64 up0 = mul_lo64(c.lo, q);
    u64 v1p = mul_hi_known(p, up0, c.lo);
    bool b;
    u64 z0 = sub_b(t.hi, v1, &b);
    m2 r;
    r.d0 = b ? z0 + p : z0;
    r.d1 = sub_mod_c(c.hi, v1p, b, p);
    return r;
}

/* REDC(x*y + k), the Montgomery product plus a constant as in m2_sqr_add. The cross sum
   x0*y1 + u0 + x1*y0 + k1 is below 2p*2^64 (at most 2p^2 - 3p + 2^64), a 129-bit value
   whose high part one conditional subtraction of p brings below p. */
WW_HD m2 m2_mul_add(m2 x, m2 y, m2 k, u64 p, u64 q, u64 np)
{
    u128 t = mul_wide_add(x.d0, y.d0, k.d0);
    u64 u0 = mul_lo64(t.lo, q);
    u64 v1 = mul_hi_known(p, u0, t.lo);
    u128 a = mul_wide_add(x.d0, y.d1, u0);
    u128 e = mul_wide_add(x.d1, y.d0, k.d1);
    bool top;
    u128 sum = add128_c(a, e, &top);
    bool big = top | (sum.hi >= p);
    u64 t1p = sum.hi + (big ? np : 0); /* minus p, modulo 2^64 */
    u64 up0 = mul_lo64(sum.lo, q);
    u64 v1p = mul_hi_known(p, up0, sum.lo);
    bool b;
    u64 z0 = sub_b(t.hi, v1, &b);
    m2 r;
    r.d0 = b ? z0 + p : z0;
    r.d1 = sub_mod_c(t1p, v1p, b, p);
    return r;
}

/* x * 2^v mod p^2 for v <= 31 (Montgomery form times an integer stays in Montgomery
   form): d0 2^v = q0 p + r0, then d1 2^v + q0 = q1 p + r1, digits (r0, r1). */
WW_HD m2 m2_mulpow2(m2 x, u32 v, u64 p, double pinv)
{
    u32 q0, q1;
    m2 r;
    r.d0 = mulpow2_add_qr(x.d0, v, 0, p, pinv, &q0);
    r.d1 = mulpow2_add_qr(x.d1, v, q0, p, pinv, &q1);
    return r;
}

/* Montgomery form back to ordinary digits (REDC, wwcpu.cpp:189-206). */
WW_HD m2 m2_get(m2 x, u64 p, u64 q)
{
    u64 u0 = mul_lo64(q, x.d0);
    u64 v1 = mul_hi_known(p, u0, x.d0);
    u64 tp = x.d1 + u0;
    u64 up0 = mul_lo64(q, tp);
    u64 t1p = tp < x.d1 ? 1 : 0;
    u64 v1p = mul_hi_known(p, up0, tp);
    bool b = v1 != 0;
    m2 r;
    r.d0 = b ? p - v1 : 0;
    r.d1 = sub_mod_c(t1p, v1p, b, p);
    return r;
}

/* x mod 5 using 2^32 = 1 (mod 5): fold the high word twice; the second fold cannot
   carry because the first sum is at most 2^33 - 2. */
WW_HD u32 mod5(u64 x)
{
    u64 s = (x & 0xffffffffu) + (x >> 32);
    u32 t = (u32)s + (u32)(s >> 32);
    return t % 5u;
}

/* x / 5 mod p for x < p: (x + j*p) / 5 with j = -x * p^-1 mod 5, computed as an exact
   division by multiplication with 5^-1 mod 2^64. Returns j through *jout. */
WW_HD u64 div5_digit(u64 x, u64 p, u32 pinv5, u32 *jout)
{
    u32 j = ((5u - mod5(x)) * pinv5) % 5u;
    *jout = j;
    /* x + j*p < 5*2^64; the quotient is < p < 2^64 so its low 64 bits suffice */
    return (x + (u64)j * p) * 0xCCCCCCCCCCCCCCCDull;
}

/* x / 5 mod p^2 in digits: y0 = x0/5 (mod p) with 5*y0 = x0 + j*p, then
   y1 = (x1 - j)/5 (mod p). */
WW_HD m2 m2_div5(m2 x, u64 p)
{
    u32 k = mod5(p);
    u32 pinv5 = (k * k * k) % 5u; /* k^-1 = k^3 mod 5 */
    u32 j0, j1;
    m2 r;
    r.d0 = div5_digit(x.d0, p, pinv5, &j0);
    u64 x1 = sub_mod(x.d1, (u64)j0, p);
    r.d1 = div5_digit(x1, p, pinv5, &j1);
    return r;
}

/* R^2 mod p^2 (R = 2^64), the Montgomery form of 2^64: one times 2^31, 2^31 and 2^2. */
WW_HD m2 m2_r2(m2 one, u64 p, double pinv)
{
    return m2_mulpow2(m2_mulpow2(m2_mulpow2(one, 31, p, pinv), 31, p, pinv), 2, p, pinv);
}

/* 2^((p-1)/2) mod p^2 as digits (s0, s1) (wieferich.cl:219-245, wwcpu.cpp:253-284), with
   r2 = R^2 mod p^2. The exponent's six bits from `top` (any bound >= its leading bit and
   >= 5, e.g. common to a wave) give K < 64 and the start 2^K as one digit-times-R^2
   product. The remaining top - 5 bits go in windows of Montgomery squares followed by one
   exact multiplication by 2^v (v < 32) instead of a square and a conditional doubling per
   bit: first (top - 5) mod 5 bits if that is not 0, then windows of 5. */
WW_HD m2 wieferich_residue_w(u64 p, u64 q, m2 r2, double pinv, int top)
{
    const int W = 5;
    u64 e = p >> 1;
    int pos = top - 5; /* bits below the start */
    u64 np = 0 - p;
    WW_LOOP_LOCAL(np); /* opaque, so that adding it is not turned back into subtracting p */
    m2 a = m2_mul_digit(1ull << (u32)(e >> pos), r2, p, q);
    while (pos > 0) {
        const int w = pos % W ? pos % W : W; /* uniform: the short window comes first */
        pos -= w;
        for (int k = 0; k < w; ++k) a = m2_sqr(a, p, q, np);
        a = m2_mulpow2(a, (u32)(e >> pos) & ((1u << w) - 1u), p, pinv);
    }
    return m2_get(a, p, q);
}

WW_HD m2 wieferich_residue(u64 p, u64 q, m2 one)
{
    const double pinv = 1.0 / (double)p;
    return wieferich_residue_w(p, q, m2_r2(one, p, pinv), pinv, 63 - clz64(p >> 1));
}

/* V_k(3, 1) = L_2k for k = 0..46: V_0 = 2, V_1 = 3, V_k+1 = 3 V_k - V_k-1. V_46 = L_92 is
   the last below 2^64. */
constexpr u32 kLucasV31Max = 46;
constexpr u64 kLucasV31[kLucasV31Max + 1] = {
    2ull, 3ull, 7ull, 18ull, 47ull, 123ull, 322ull, 843ull, 2207ull, 5778ull, 15127ull,
    39603ull, 103682ull, 271443ull, 710647ull, 1860498ull, 4870847ull, 12752043ull,
    33385282ull, 87403803ull, 228826127ull, 599074578ull, 1568397607ull, 4106118243ull,
    10749957122ull, 28143753123ull, 73681302247ull, 192900153618ull, 505019158607ull,
    1322157322203ull, 3461452808002ull, 9062201101803ull, 23725150497407ull,
    62113250390418ull, 162614600673847ull, 425730551631123ull, 1114577054219522ull,
    2918000611027443ull, 7639424778862807ull, 20000273725560978ull, 52361396397820127ull,
    137083915467899403ull, 358890350005878082ull, 939587134549734843ull,
    2459871053643326447ull, 6440026026380244498ull, 16860207025497407047ull};

/* The Wall-Sun-Sun ladder's n = (p - (p/5)) / 2. */
WW_HD u64 wallsunsun_n(u64 p)
{
    u32 pm5 = mod5(p);
    return ((pm5 == 2 || pm5 == 3) ? p + 1 : p - 1) >> 1;
}

/* How many leading ladder steps (bits top .. top-skip+1 of n) kLucasV31 can stand for:
   6 when those six bits give k <= 45 (so V_k+1 is in the table), else 5 (k <= 31).
   Every p >= 127 prime to 5 has n >= 64, so top >= 6 and the ladder keeps a bit. */
WW_HD int wallsunsun_skip(u64 n, int top)
{
    return top >= 6 && (n >> (top - 5)) < kLucasV31Max ? 6 : 5;
}

/* F_{p-(p/5)} mod p^2 as digits (c20, c21) (wallsunsun.cl:242-300, wwcpu.cpp:311-372).
   m = p - (p/5) is even, and F_{2n} = U_n(3, 1), the Lucas sequence with P = 3, Q = 1
   (its roots are the squares of the golden ratio and its conjugate). Its companion V
   has constant-coefficient doubling formulas,
     V_2k = V_k^2 - 2,   V_2k+1 = V_k V_k+1 - 3,
   so a ladder on (V_k, V_k+1) over the bits of n = m/2 costs one square and one product
   per bit with no sign-dependent constants, and 5 U_n = 2 V_{n+1} - 3 V_n at the end.
   The pair lives in two fixed slots (x, y): the product is symmetric and goes to x, the
   square of the operand the bit picks goes to y, so after bit i slot x holds V_k exactly
   when bit i is 1, and the square's operand is y iff bits i and i+1 of n are equal.
   Leading zero bits leave the ladder's start (V_0, V_1) = (2, 3) unchanged, so any top in
   [leading bit of n, 63] (e.g. common to a wave) gives the same result. The first `skip` steps (5 or 6; skip = wallsunsun_skip(n, top) or 5, or the smaller of
   several lanes' values) are read from kLucasV31 instead: k = the bits of n above
   top - skip, and V_k, V_k+1 (below 2^64) enter Montgomery form as products with R^2,
   two products in place of 2 skip. */
WW_HD m2 wallsunsun_residue_w(u64 p, u64 q, m2 r2, int top, int skip)
{
    u64 n = wallsunsun_n(p);
    u64 same = ~(n ^ (n >> 1)); /* bit i: bits i and i+1 of n are equal */
    m2 r2x2 = m2_add(r2, r2, p);
    m2 minus2 = m2_neg(r2x2, p);                /* -2 R^2 */
    m2 minus3 = m2_neg(m2_add(r2x2, r2, p), p); /* -3 R^2 */
    u64 np = 0 - p;
    WW_LOOP_LOCAL(np); /* opaque, so that adding it is not turned back into subtracting p */
    /* bits top .. start+1 done, k = n >> (start+1): slot x holds V_k iff k is odd */
    const int start = top - skip; /* next bit of n */
    const u32 k = (u32)(n >> (start + 1));
    const m2 v0 = m2_mul_digit(kLucasV31[k], r2, p, q);
    const m2 v1 = m2_mul_digit(kLucasV31[k + 1], r2, p, q);
    const bool kodd = k & 1;
    m2 x, y;
    x.d0 = kodd ? v0.d0 : v1.d0;
    x.d1 = kodd ? v0.d1 : v1.d1;
    y.d0 = kodd ? v1.d0 : v0.d0;
    y.d1 = kodd ? v1.d1 : v0.d1;
    /* The bits of `same` are read one 32-bit word at a time, high word first when start
       >= 32: the word is doubled each step and the carry out is the bit (one
       v_add_co_u32 on the device gives both). */
    for (int half = start >= 32 ? 1 : 0; half >= 0; --half) {
        const int lead = half ? start - 32 : (start >= 32 ? 31 : start); /* first bit, in the word */
        u32 w = (half ? (u32)(same >> 32) : (u32)same) << (31 - lead);
        for (int j = lead; j >= 0; --j) {
            WW_LOOP_LOCAL(p);
            bool use_y = __builtin_add_overflow(w, w, &w);
            m2 sin;
            sin.d0 = use_y ? y.d0 : x.d0;
            sin.d1 = use_y ? y.d1 : x.d1;
            x = m2_mul_add(x, y, minus3, p, q, np); /* V_{2k+1} = V_k V_{k+1} - 3 */
            y = m2_sqr_add(sin, minus2, p, q, np);  /* V_{2k} or V_{2k+2} = V^2 - 2 */
        }
    }
    /* slot x holds V_n iff bit 0 of n is 1 */
    bool odd = n & 1;
    m2 vk, vk1;
    vk.d0 = odd ? x.d0 : y.d0;
    vk.d1 = odd ? x.d1 : y.d1;
    vk1.d0 = odd ? y.d0 : x.d0;
    vk1.d1 = odd ? y.d1 : x.d1;
    m2 two_v1 = m2_add(vk1, vk1, p);
    m2 three_v0 = m2_add(m2_add(vk, vk, p), vk, p);
    m2 f5 = m2_sub(two_v1, three_v0, p);
    return m2_get(m2_div5(f5, p), p, q);
}

/* Bit index of the leading bit of the Wall-Sun-Sun ladder's n for this p. */
WW_HD int wallsunsun_top(u64 p)
{
    return 63 - clz64(wallsunsun_n(p));
}

WW_HD m2 wallsunsun_residue(u64 p, u64 q, m2 one)
{
    const int top = wallsunsun_top(p);
    return wallsunsun_residue_w(p, q, m2_r2(one, p, 1.0 / (double)p), top,
                                wallsunsun_skip(wallsunsun_n(p), top));
}

} // namespace ww

#endif
