/*
 * ww_bucket.h -- arithmetic of the bucket sieve: the sieving primes above the wide
 * pre-sieve, up to the square root of a range's end.
 *
 * Compiles for the device (the fill kernel, ww_bucket.hip) and for the host (the
 * verification entry point ww_cpu_bucket_hits), so the CPU tests pin the very code the
 * kernel runs.
 *
 * For a sieving prime l and the odd numbers G0 + 2b of a range (G0 odd), the multiples
 * n = l*k are walked only for k prime to 30030 = 2*3*5*7*11*13: n is odd and prime to
 * 3..13 exactly when k is, and any other n is no candidate (every candidate is odd and
 * prime to 3..113). That skips 9255 of every 15015 odd multiples; those with 17, 19 or 23
 * dividing k, and those with a prime factor from 29 to 53 dividing k, are stepped over
 * without a visit (k mod 215441 = 17 x 19 x 23 x 29, k mod 47027 = 31 x 37 x 41 and
 * k mod 107113 = 43 x 47 x 53 are carried and looked up in three bitmaps), which leaves
 * 0.272 of them. Each visited n
 * also gets its base-2 pseudoprime test modulo l. A composite n = l*k that is a base-2 Fermat
 * pseudoprime (every strong one is) has 2^(n-1) = 1 (mod l); such an n is not dropped but
 * flagged for the strong PRP test. By Fermat, 2^(n-1) = 2^(lk-1) = (2^(l-1))^k 2^(k-1) =
 * 2^(k-1) (mod l), so the test is ord_l(2) | k - 1: the walk carries c = (k - 1) mod ord_l(2)
 * and adds each wheel gap to it, and the orders of all sieving primes come from a table
 * built once per context (order2_from_factors, from the factors of l - 1).
 */
#ifndef WW_BUCKET_H
#define WW_BUCKET_H

#include "ww_arith.h"

namespace ww {

/* Cofactor wheel modulo 30030: its 5760 residues prime to 30030, ascending, the gaps
   between them, and the rank structure that gives the index of the first residue >= r
   (a bitmap of the residues and the count of those below each 32-bit word). The gaps are
   stored twice over: a walk covers fewer than 30030 consecutive cofactors (kWalkMaxK), so
   it passes at most 5760 residues and its index, from below 5760, stays below 2 x 5760
   without wrapping. Beside them, the bitmaps of the residues of k mod 215441 =
   17 x 19 x 23 x 29, k mod 47027 = 31 x 37 x 41 and k mod 107113 = 43 x 47 x 53 prime to
   their moduli: a cofactor is visited when all three bits are set. */
constexpr u32 kWheelMod = 30030;
constexpr u32 kWheelRes = 5760;
constexpr u32 kWheelWords = (kWheelMod + 31) / 32;
constexpr u32 kVisitModA = 17 * 19 * 23 * 29; /* 215441 */
constexpr u32 kVisitModB = 31 * 37 * 41;      /* 47027 */
constexpr u32 kVisitModC = 43 * 47 * 53;      /* 107113 */
constexpr u32 kVisitWordsA = (kVisitModA + 31) / 32;
constexpr u32 kVisitWordsB = (kVisitModB + 31) / 32;
constexpr u32 kVisitWordsC = (kVisitModC + 31) / 32;
/* Most cofactors k in one walk: the window [kmin, kmax) is at most this long. */
constexpr u64 kWalkMaxK = kWheelMod - 1;

struct WheelTables {
    uint16_t res[kWheelRes];       /* the residues */
    uint8_t gap[2 * kWheelRes];    /* res[i + 1] - res[i], and 30030 + res[0] - res[5759];
                                      entry i + 5760 = entry i */
    uint32_t bits[kWheelWords];    /* bit r % 32 of word r / 32: r is a residue */
    uint16_t rank[kWheelWords];    /* residues below 32 w */
    uint32_t visit_a[kVisitWordsA]; /* bit m % 32 of word m / 32: m prime to 215441 */
    uint32_t visit_b[kVisitWordsB]; /* the same for 47027 */
    uint32_t visit_c[kVisitWordsC]; /* and for 107113 */
};

constexpr WheelTables make_wheel()
{
    WheelTables t{};
    u32 n = 0;
    for (u32 r = 0; r < kWheelMod; ++r)
        if (r % 2 && r % 3 && r % 5 && r % 7 && r % 11 && r % 13) {
            t.res[n++] = (uint16_t)r;
            t.bits[r / 32] |= 1u << (r % 32);
        }
    for (u32 i = 0; i < kWheelRes; ++i) {
        const u32 nx = i + 1 < kWheelRes ? t.res[i + 1] : kWheelMod + t.res[0];
        t.gap[i] = t.gap[i + kWheelRes] = (uint8_t)(nx - t.res[i]);
    }
    u32 below = 0;
    for (u32 w = 0; w < kWheelWords; ++w) {
        t.rank[w] = (uint16_t)below;
        for (u32 b = 0; b < 32; ++b) below += (t.bits[w] >> b) & 1u;
    }
    /* the visit bitmaps as sieves: every bit set, then the multiples of each factor
       cleared (few enough steps for a constant expression) */
    for (u32 w = 0; w < kVisitWordsA; ++w) t.visit_a[w] = ~0u;
    for (u32 q : {17u, 19u, 23u, 29u})
        for (u32 m = 0; m < kVisitModA; m += q) t.visit_a[m / 32] &= ~(1u << (m % 32));
    for (u32 w = 0; w < kVisitWordsB; ++w) t.visit_b[w] = ~0u;
    for (u32 q : {31u, 37u, 41u})
        for (u32 m = 0; m < kVisitModB; m += q) t.visit_b[m / 32] &= ~(1u << (m % 32));
    for (u32 w = 0; w < kVisitWordsC; ++w) t.visit_c[w] = ~0u;
    for (u32 q : {43u, 47u, 53u})
        for (u32 m = 0; m < kVisitModC; m += q) t.visit_c[m / 32] &= ~(1u << (m % 32));
    return t;
}

/* The wheel's largest gap: bucket_advance's single subtractions mod ord (ord >= 22) and
   mod the two visit moduli rely on it. */
constexpr u32 kWheelMaxGap = 22;

constexpr u32 wheel_max_gap()
{
    constexpr WheelTables t = make_wheel();
    u32 g = 0;
    for (u32 i = 0; i < kWheelRes; ++i) g = t.gap[i] > g ? t.gap[i] : g;
    return g;
}
static_assert(wheel_max_gap() == kWheelMaxGap, "the wheel's largest gap");

/* Index of the first residue >= r (r < 30030), kWheelRes when there is none. */
WW_HD u32 wheel_index(const uint32_t *bits, const uint16_t *rank, u32 r)
{
    return rank[r >> 5] + (u32)__builtin_popcount(bits[r >> 5] & ((1u << (r & 31)) - 1u));
}

/* x = q d + r for x < 2^64 and d < 2^32 with x/d < 2^44 (any x when d >= 2^20), with
   dinv = 1/d rounded to double (or inv_below(d)). The double quotient is within 2^-49 of
   x/d relatively, and x/d < 2^44, so it is off by less than 2^-5: truncated it is the true quotient or one away from it, and
   the signed remainder x - q d (exact mod 2^64, its true value lies in (-d, 2d)) decides
   the one correction. */
WW_HD void divmod_u64_u32(u64 x, u32 d, double dinv, u64 *q, u32 *r)
{
    u64 qe = (u64)((double)x * dinv);
    int64_t rs = (int64_t)(x - qe * (u64)d);
    if (rs < 0) {
        rs += d;
        --qe;
    } else if (rs >= (int64_t)d) {
        rs -= d;
        ++qe;
    }
    *q = qe;
    *r = (u32)rs;
}

/* 1/l rounded to double and nudged down by 2^-50 relatively: every quotient estimate
   below is then at most the true quotient and, for quotients below 2^32, at most one
   below it (its relative error stays within 2^-49). */
WW_HD double inv_below(u32 l) { return (1.0 / (double)l) * (1.0 - 0x1p-50); }

/* x 2^v mod l for x < l < 2^32, v <= 32 and x 2^v < 2^63, with linv = inv_below(l). The
   quotient is below 2^v <= 2^32, and the estimate from the exact double x 2^v is the true
   quotient or one less, so the remainder x 2^v - q l lies in [0, 2l) and one conditional
   subtraction finishes it. */
WW_HD u32 mulpow2_u32(u32 x, u32 v, u32 l, double linv)
{
    const u64 t = (u64)x << v;
    const u32 q = (u32)(ldexp((double)x, (int)v) * linv);
    const u64 r = t - (u64)q * l;
    return (u32)r - (r >= (u64)l ? l : 0u);
}

/* -l^-1 mod 2^32 for odd l: l * l = 1 (mod 8), and each Newton step y(2 - l y) doubles
   the correct low bits (3, 6, 12, 24, 48). */
WW_HD u32 neg_inv_u32(u32 l)
{
    u32 y = l;
    y *= 2u - l * y;
    y *= 2u - l * y;
    y *= 2u - l * y;
    y *= 2u - l * y;
    return 0u - y;
}

/* Montgomery product a b 2^-32 mod l for a, b < l < 2^32 odd, linv = -l^-1 mod 2^32.
   t + m l = 0 (mod 2^32), so (t + m l) / 2^32 = hi(t) + hi(m l) + carry, and the low words
   add up to 0 or 2^32: the carry is 1 exactly when lo(t) != 0. hi(t) + carry stays below
   2^32 (t < l^2), and the full sum, below l^2 / 2^32 + l < 2l, may pass 2^32 once: then, or
   when it is at least l, one subtraction of l finishes it (mod 2^32 arithmetic). */
WW_HD u32 mont_mul_u32(u32 a, u32 b, u32 l, u32 linv)
{
    const u64 t = (u64)a * b;
    const u32 tl = (u32)t, th = (u32)(t >> 32);
    const u32 m = tl * linv;
    const u32 mh = (u32)(((u64)m * l) >> 32);
    const u32 s = th + (tl != 0u ? 1u : 0u);
    const u32 u = s + mh;
    return (u < s || u >= l) ? u - l : u;
}

/* 2^e mod l for e < 2^32 and 2^20 <= l < 2^32 odd (linv = inv_below(l)): Montgomery
   squarings over 4-bit windows of e from the top, each window's 2^w entered by one exact
   multiplication. */
WW_HD u32 pow2_mod_u32(u32 e, u32 l, double linv)
{
    const u32 li = neg_inv_u32(l);
    const u32 r1 = mulpow2_u32(1u, 32u, l, linv); /* 2^32 mod l: Montgomery 1 */
    u32 x = mulpow2_u32(r1, e >> 28, l, linv);
    for (int s = 24; s >= 0; s -= 4) {
        x = mont_mul_u32(x, x, l, li);
        x = mont_mul_u32(x, x, l, li);
        x = mont_mul_u32(x, x, l, li);
        x = mont_mul_u32(x, x, l, li);
        x = mulpow2_u32(x, (e >> s) & 15u, l, linv);
    }
    return mont_mul_u32(x, 1u, l, li);
}

/* The odd primes below 2^16 with the data of an exact divisibility test: for odd q and any
   32-bit m, q | m iff m qinv mod 2^32 <= qlim, where qinv = q^-1 mod 2^32 and
   qlim = floor((2^32 - 1) / q) (multiplication by qinv permutes the residues, and maps the
   multiples m = q j onto j <= qlim); then m / q = m qinv mod 2^32. */
struct SmallDiv {
    u32 q, qinv, qlim, pad;
};
constexpr u32 kSmallOddPrimes = 6541; /* odd primes below 2^16 */

WW_HD bool small_divides(u32 m, const SmallDiv &t) { return m * t.qinv <= t.qlim; }

/* At most 9 distinct odd primes below 2^16 divide a 32-bit number (3 x 5 x ... x 29 is
   already above 2^31), and at most one prime above 2^16: ten factor slots. */
constexpr u32 kOrderFactors = 10;

/* ord_l(2) for a prime 2^20 < l < 2^32, from the distinct odd prime factors f[0..nf) of
   l - 1 (linv = inv_below(l)). Starting from o = l - 1, a multiple of the order, each odd
   prime q is divided out of o while 2^(o/q) = 1: since the order divides o, that holds
   exactly while q divides o more often than the order, so q ends with the order's exponent,
   whatever the order of the primes. That leaves o = 2^s x (odd part of the order); the
   order's power of two is the number of squarings that take 2^(odd part) to 1. The slots
   are visited in an unrolled loop, so the factors stay in registers on the device. */
WW_HD u32 order2_from_factors(u32 l, double linv, const u32 *f, u32 nf)
{
    u32 o = l - 1;
#pragma unroll
    for (u32 j = 0; j < kOrderFactors; ++j) {
        if (j >= nf) continue;
        const u32 q = f[j];
        while (o % q == 0 && pow2_mod_u32(o / q, l, linv) == 1u) o /= q;
    }
    const u32 s = (u32)__builtin_ctz(l - 1);
    const u32 odd = o >> s;
    u32 y = pow2_mod_u32(odd, l, linv), t = 0;
    while (y != 1u) { /* at most s squarings: 2^(l-1) = 1 */
        u64 qq;
        divmod_u64_u32((u64)y * y, l, linv, &qq, &y);
        ++t;
    }
    return odd << t;
}

/* The distinct odd prime factors of l - 1 (l odd, 2^20 < l < 2^32) by trial division with
   the odd primes below 2^16 (tab[0..ntab)), stopping once q^2 exceeds the unfactored part
   m, which is then 1 or a prime (also when the table runs out: every composite below 2^32
   has a factor below 2^16). Returns the number of factors written to f. */
WW_HD u32 odd_factors_u32(u32 l, const SmallDiv *tab, u32 ntab, u32 *f)
{
    u32 m = (l - 1) >> __builtin_ctz(l - 1), nf = 0;
    for (u32 j = 0; j < ntab && tab[j].q * tab[j].q <= m; ++j) {
        if (!small_divides(m, tab[j])) continue;
        f[nf++] = tab[j].q;
        do m *= tab[j].qinv; while (small_divides(m, tab[j]));
    }
    if (m > 1u) f[nf++] = m;
    return nf;
}

/* ord_l(2) for a prime 2^20 < l < 2^32 (the host's way; the table kernel runs the same
   factor search as one wave-wide loop). Every such order is at least 22: 2^ord > l > 2^20
   needs ord >= 21, and 2^21 - 1 = 7^2 x 127 x 337 has no prime factor above 2^20. */
WW_HD u32 order2_u32(u32 l, const SmallDiv *tab, u32 ntab)
{
    u32 f[kOrderFactors];
    const u32 nf = odd_factors_u32(l, tab, ntab, f);
    return order2_from_factors(l, inv_below(l), f, nf);
}

/* The walk of the sieving prime l (2^20 < l < 2^32, ord = ord_l(2)) over the odd numbers
   G0 + 2b (G0 odd): the multiples n = l k with k prime to 30030, from the first one >= G0
   on, for at most kWalkMaxK consecutive cofactors k. d = n - G0 is the current multiple's
   distance, twice its bit (the true distance is below 22 l at the start, so mod-2^64
   arithmetic gives it even where n passes 2^64), idx the wheel index of k, c =
   (k - 1) mod ord, ma = k mod 215441, mb = k mod 47027 and mc = k mod 107113: a multiple
   whose cofactor has a factor from 17 to 53 is no candidate either, so it is stepped over
   but not visited (bucket_visit), which leaves 16/17 x 18/19 x ... x 52/53 of the wheel's
   multiples. */
struct BucketWalk {
    u64 d;
    u32 idx;
    u32 c;
    u32 ma, mb, mc;
};

WW_HD BucketWalk bucket_start(u64 G0, u32 l, double linv, u32 ord, const uint16_t *res,
                              const uint32_t *bits, const uint16_t *rank)
{
    /* linv = inv_below(l) */
    u64 q;
    u32 r;
    divmod_u64_u32(G0, l, linv, &q, &r);
    const u64 kmin = q + (r != 0u ? 1u : 0u); /* ceil(G0 / l), below 2^44 + 1 */
    u64 kq = kmin / kWheelMod;
    u32 idx = wheel_index(bits, rank, (u32)(kmin - kq * kWheelMod));
    if (idx == kWheelRes) {
        idx = 0;
        ++kq;
    }
    const u64 k = kq * kWheelMod + res[idx];
    BucketWalk w;
    w.d = (u64)l * k - G0;
    w.idx = idx;
    /* k - 1 < 2^44 + 2^5 and ord >= 22: the quotient is below 2^44 */
    u64 q2;
    divmod_u64_u32(k - 1, ord, inv_below(ord), &q2, &w.c);
    /* k mod 215441, 47027 and 107113 (k < 2^45; 30030 is below the moduli) */
    w.ma = (u32)(((u64)((u32)kq % kVisitModA) * kWheelMod + res[idx]) % kVisitModA);
    w.mb = ((u32)kq % kVisitModB * kWheelMod + res[idx]) % kVisitModB;
    w.mc = (u32)(((u64)((u32)kq % kVisitModC) * kWheelMod + res[idx]) % kVisitModC);
    return w;
}

/* Whether the walk's current multiple is a candidate position (its cofactor prime to 17
   .. 53 by the bitmaps; prime to 30030 by the wheel). */
WW_HD bool bucket_visit(const BucketWalk &w, const uint32_t *visit_a, const uint32_t *visit_b,
                        const uint32_t *visit_c)
{
    return (visit_a[w.ma >> 5] >> (w.ma & 31u)) & (visit_b[w.mb >> 5] >> (w.mb & 31u)) &
           (visit_c[w.mc >> 5] >> (w.mc & 31u)) & 1u;
}

/* The next multiple, the wheel gap g = gap[idx] ahead: k -> k + g, d -> d + g l,
   c -> (c + g) mod ord, ma, mb and mc -> (ma + g, mb + g, mc + g) mod their moduli, each
   with one subtraction (g <= 22 <= ord),
   idx -> idx + 1 (the gaps are stored twice, so no wrap within a walk). */
WW_HD void bucket_advance(BucketWalk &w, u32 l, u32 ord, u32 g)
{
    w.d += (u64)g * l;
    w.c += g;
    w.c = w.c >= ord ? w.c - ord : w.c;
    w.ma += g;
    w.ma = w.ma >= kVisitModA ? w.ma - kVisitModA : w.ma;
    w.mb += g;
    w.mb = w.mb >= kVisitModB ? w.mb - kVisitModB : w.mb;
    w.mc += g;
    w.mc = w.mc >= kVisitModC ? w.mc - kVisitModC : w.mc;
    ++w.idx;
}

WW_HD void bucket_step(BucketWalk &w, u32 l, u32 ord, const uint8_t *gap)
{
    bucket_advance(w, l, ord, gap[w.idx]);
}

/* Bucket entries: the odd-number index of the hit in its 2^20-bit chunk, and the top bit
   set when the hit is in the prime's pseudoprime class (c = 0). */
constexpr u32 kBucketChunkBits = 20;
constexpr u32 kBucketRescue = 0x80000000u;

/* The kSmallOddPrimes entries of SmallDiv, built once on the host (ww_cpu.cpp). */
const SmallDiv *small_div_table();

} // namespace ww

#endif
