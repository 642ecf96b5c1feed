/*
 * ww_bucket.h -- arithmetic of the bucket sieve: the sieving primes above the wide
 * pre-sieve, up to the square root of a range's end.
 *
 * Compiles for the device (the fill kernel, ww_bucket.hip) and for the host (the
 * verification entry point ww_cpu_bucket_hits), so the CPU tests pin the very code the
 * kernel runs.
 *
 * For a sieving prime l and the odd numbers G0 + 2b of a range (G0 odd), the multiples
 * n = l*k are visited only for k prime to 30030 = 2*3*5*7*11*13: n is odd and prime to
 * 3..13 exactly when k is, and any other n is no candidate (every candidate is odd and
 * prime to 3..113). That skips 9255 of every 15015 odd multiples. Each visited n also
 * gets its base-2 pseudoprime test modulo l: x = 2^(n-1) mod l. A composite n = l*k that is a base-2
 * Fermat pseudoprime (every strong one is) has x = 1, i.e. ord_l(2) | n - 1; such an n is
 * not dropped but flagged for the strong PRP test. By Fermat, 2^(n-1) = 2^(lk-1) =
 * (2^(l-1))^k 2^(k-1) = 2^(k-1) (mod l), so x starts as 2^(k0-1) mod l and each wheel
 * step k -> k + g multiplies it by 2^g.
 */
#ifndef WW_BUCKET_H
#define WW_BUCKET_H

#include "ww_arith.h"

namespace ww {

/* Cofactor wheel modulo 30030: its 5760 residues prime to 30030, ascending, the gaps
   between them, and the rank structure that gives the index of the first residue >= r
   (a bitmap of the residues and the count of those below each 32-bit word). */
constexpr u32 kWheelMod = 30030;
constexpr u32 kWheelRes = 5760;
constexpr u32 kWheelWords = (kWheelMod + 31) / 32;

struct WheelTables {
    uint16_t res[kWheelRes];     /* the residues */
    uint8_t gap[kWheelRes];      /* res[i + 1] - res[i], and 30030 + res[0] - res[5759] */
    uint32_t bits[kWheelWords];  /* bit r % 32 of word r / 32: r is a residue */
    uint16_t rank[kWheelWords];  /* residues below 32 w */
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
        t.gap[i] = (uint8_t)(nx - t.res[i]);
    }
    u32 below = 0;
    for (u32 w = 0; w < kWheelWords; ++w) {
        t.rank[w] = (uint16_t)below;
        for (u32 b = 0; b < 32; ++b) below += (t.bits[w] >> b) & 1u;
    }
    return t;
}

/* The wheel's largest gap: 2^22 x < 2^54 keeps every x * 2^gap in mulpow2_u32's range. */
constexpr u32 kWheelMaxGap = 22;

/* Index of the first residue >= r (r < 30030), kWheelRes when there is none. */
WW_HD u32 wheel_index(const uint32_t *bits, const uint16_t *rank, u32 r)
{
    return rank[r >> 5] + (u32)__builtin_popcount(bits[r >> 5] & ((1u << (r & 31)) - 1u));
}

/* x = q d + r for x < 2^64 and 2^20 <= d < 2^32, with dinv = 1/d rounded to double (or
   inv_below(d)). The double quotient is within 2^-49 of x/d relatively, and x/d < 2^44, so
   it is off by less than 2^-5: truncated it is the true quotient or one away from it, and
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

/* 2^e mod l for e < 2^32 and 2^20 <= l < 2^32 prime (linv = inv_below(l)): Montgomery
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

/* Where the bucket sieve starts for the sieving prime l (2^20 < l < 2^32) over the odd
   numbers G0 + 2b (G0 odd): the first multiple n = l k >= G0 with k prime to 30030, as its
   bit b (the true difference n - G0 is below 22 l, so mod-2^64 arithmetic gives it even
   where n passes 2^64), the wheel index of k, and x = 2^(n-1) mod l. */
struct BucketStart {
    u64 b;
    u32 idx;
    u32 x;
};

WW_HD BucketStart bucket_start(u64 G0, u32 l, double linv, const uint16_t *res,
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
    BucketStart s;
    s.b = ((u64)l * k - G0) >> 1;
    s.idx = idx;
    /* x = 2^(k-1) mod l, the exponent reduced mod l - 1 */
    u64 q2;
    u32 e;
    divmod_u64_u32(k - 1, l - 1, inv_below(l - 1), &q2, &e);
    s.x = pow2_mod_u32(e, l, linv);
    return s;
}

/* The next visited multiple: k -> k + gap, b -> b + gap l / 2, x -> x 2^gap. */
WW_HD void bucket_step(u64 &b, u32 &idx, u32 &x, u32 l, double linv, const uint8_t *gap)
{
    const u32 g = gap[idx];
    b += (u64)(g >> 1) * l;
    x = mulpow2_u32(x, g, l, linv);
    idx = idx + 1 == kWheelRes ? 0u : idx + 1;
}

/* Bucket entries: the odd-number index of the hit in its 2^20-bit chunk, and the top bit
   set when the hit is in the prime's pseudoprime class (x = 1). */
constexpr u32 kBucketChunkBits = 20;
constexpr u32 kBucketRescue = 0x80000000u;

} // namespace ww

#endif
