/*
 * ww_device.h -- device pieces shared by the plain and the persistent kernels: the
 * small-prime word sieve, result classification and record append.
 */
#ifndef WW_DEVICE_H
#define WW_DEVICE_H

#include "ww_arith.h"
#include "ww_kernels.h"

namespace ww {

/* Bits 0, L, 2L, ... of a 64-bit word: one lane's word holds odd numbers
   w0, w0+2, ..., w0+126, so multiples of L recur every L bits. */
__device__ constexpr u64 stride_pattern(u32 L)
{
    u64 m = 0;
    for (u32 i = 0; i < 64; i += L) m |= 1ull << i;
    return m;
}

/* Marks the bits of `comp` whose odd number w0 + 2i is divisible by L, for a lane word at
   w0 = group + 2 half_off (half_off < 2^15) and gneg = (-group / 2) mod L (wave-uniform).
   L | w0 + 2i exactly when i = (gneg - half_off) mod L. x = gneg + C - half_off with C a
   multiple of L above 2^15 stays below 2^16, where the quotient mulhi(x, ceil(2^32 / L))
   is exact (its error is below x / 2^32 < 1/L), so the residue costs three VALU
   instructions per prime and lane. */
template <u32 L>
__device__ __forceinline__ void sieve_prime(u64 &comp, u32 gneg, u32 half_off)
{
    constexpr u64 pat = stride_pattern(L);
    constexpr u32 C = L * ((1u << 15) / L + 1);
    constexpr u32 M = (u32)(((1ull << 32) + L - 1) / L);
    const u32 x = gneg + C - half_off;
    const u32 i0 = x - __umulhi(x, M) * L; /* first i with L | w0 + 2i */
    if (L < 64) {
        comp |= pat << i0;
    } else {
        comp |= (i0 < 64) ? (1ull << i0) : 0ull;
    }
}

/* The odd primes 3..113: together with oddness, the candidate filter of the reference's
   wheel + bit tables (getsegprimes.cl:167-278) and of its CPU loop (main.cpp:653-682). */
#define WW_SIEVE_PRIMES(X)                                                            \
    X(3) X(5) X(7) X(11) X(13) X(17) X(19) X(23) X(29) X(31) X(37) X(41) X(43) X(47) \
    X(53) X(59) X(61) X(67) X(71) X(73) X(79) X(83) X(89) X(97) X(101) X(103) X(107) \
    X(109) X(113)

/* Candidate bits of the word of odd numbers w0 + 2i (i < 64), w0 = group + lane_off
   (lane_off even, below 2^16): no factor in 3..113, and w0 + 2i < hi. Words that wrapped
   past 2^64 (w0 < group) yield nothing. `group` is uniform across the wave, so its
   residues are scalar work. */
__device__ __forceinline__ u64 sieve_word(u64 group, u32 lane_off, u64 hi)
{
    const u64 w0 = group + lane_off;
    const u32 half_off = lane_off >> 1;
    u64 comp = 0;
    /* gneg = (-group) * 2^-1 mod L, with 2^-1 = (L + 1) / 2 */
#define WW_SIEVE_ONE(L)                                                                  \
    sieve_prime<L>(comp, (u32)(((L - group % L) % L) * ((L + 1) / 2) % L), half_off);
    WW_SIEVE_PRIMES(WW_SIEVE_ONE)
#undef WW_SIEVE_ONE
    u64 valid = 0;
    if (w0 >= group && w0 < hi) {
        u64 left = hi - w0;       /* >= 1 */
        u64 nv = (left + 1) >> 1; /* odd numbers in [w0, hi) */
        valid = nv >= 64 ? ~0ull : ((1ull << nv) - 1);
    }
    return ~comp & valid;
}

/* ---------------------------------------------------------------------------
 * Deep sieve by the primes 127 <= l <= B over one wave's 4096 odd numbers, with the
 * exact rescue of base-2 pseudoprimes.
 *
 * A composite n = l*k that is a base-2 Fermat pseudoprime (every strong one is) has
 * 2^(n-1) = 1 (mod l), so ord_l(2) | n - 1: n lies in the single class n = 0 (mod l),
 * n = 1 (mod lcm(2, ord_l(2))), i.e. n = rho (mod M) with M = l * lcm(2, ord_l(2)).
 * Marking every multiple of l and then clearing that class again therefore removes only
 * numbers that cannot be base-2 strong probable primes; l itself (a prime candidate) is
 * never marked. The strong PRP test still decides membership, so the list is unchanged.
 * ------------------------------------------------------------------------- */

/* Largest x over the 64 lanes of the wave, wave-uniform. Every lane must take part. */
__device__ __forceinline__ int wave_max(int x)
{
    for (int off = 32; off >= 1; off >>= 1) x = max(x, __shfl_xor(x, off, 64));
    return __builtin_amdgcn_readfirstlane(x);
}

/* Wave-uniform values as scalars (the compiler cannot see that a lane-0 broadcast is
   uniform, and would keep them in vector registers). */
__device__ __forceinline__ u32 uniform32(u32 v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ u64 uniform64(u64 v)
{
    return ((u64)uniform32((u32)(v >> 32)) << 32) | uniform32((u32)v);
}

/* Wave compaction of a segment's candidates: bit b of lane L's word stands for the odd
   number G + 2 i, i = 64 L + b, and wc[base ...] receives tag | i for every set bit. The
   entries are grouped by t, the trailing-zero count of n - 1 (t = 1, 2, 3, >= 4), so that
   most strong PRP rounds hold one group and their bit-by-bit tail below the wave's largest
   t stays short. n - 1 = 2 (g + i) with g = (G - 1) / 2, so t = 1 + ctz(g + i) depends only
   on (g + i) mod 8 (64 L = 0 mod 8): four bit masks shared by every lane. The lanes scan
   their four counts at once, in 16-bit fields (a segment holds at most 4096). The groups
   go in ascending t, or descending when `desc` (wave-uniform): alternating it between
   consecutive segments makes a round that mixes one segment's last entries with the next
   one's first stay within one group. Returns the number of entries written, wave-uniform.
   Called by all 64 lanes. */
__device__ __forceinline__ u32 compact_by_class(u64 bits, u64 G, u32 lane, uint16_t *wc,
                                                u32 base, u32 tag, bool desc = false)
{
    const u32 g8 = (u32)((G - 1) >> 1) & 7u;
    u64 cls[4];
    cls[0] = 0x5555555555555555ull << ((1u - g8) & 1u); /* g + i odd: t = 1 */
    cls[1] = 0x1111111111111111ull << ((2u - g8) & 3u); /* = 2 mod 4: t = 2 */
    cls[2] = 0x0101010101010101ull << ((4u - g8) & 7u); /* = 4 mod 8: t = 3 */
    cls[3] = 0x0101010101010101ull << ((0u - g8) & 7u); /* = 0 mod 8: t >= 4 */
    u64 cnt = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) cnt |= (u64)__popcll(bits & cls[k]) << (16 * k);
    u64 incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
        u64 v = __shfl_up(incl, d, 64);
        if (lane >= (u32)d) incl += v;
    }
    const u64 tot = uniform64(__shfl(incl, 63, 64));
    u32 at = base;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
        const int k = desc ? 3 - kk : kk;
        u32 pos = at + (u32)(((incl - cnt) >> (16 * k)) & 0xffffu);
        u64 b = bits & cls[k];
        while (b) {
            wc[pos++] = (uint16_t)(tag | (lane * 64 + (u32)__builtin_ctzll(b)));
            b &= b - 1;
        }
        at += (u32)((tot >> (16 * k)) & 0xffffu);
    }
    return at - base;
}

/* Base-2 strong PRP of one candidate per lane (act false: the lane only takes part in
   the wave's common loop bounds, the largest leading bit and trailing-zero count of
   n - 1, so that every lane runs the same loops). Called by all 64 lanes. */
__device__ __forceinline__ bool spsp2_wave(u64 n, bool act)
{
    const u64 e = act ? n - 1 : 2;
    const int top = wave_max(63 - clz64(e));
    const int tmax = wave_max(ctz64(e));
    bool prp = false;
    if (act) prp = spsp2_window(n, inv64(n), mont_one(n), top, tmax);
    return prp;
}

/* Deep sieve of one wave segment (odd numbers group + 2i, i < 4096): returns this
   lane's 64-bit word of positions to drop, the multiples of the deep primes minus each
   prime's base-2 pseudoprime class (which holds every strong pseudoprime that the prime
   divides, so no list member is dropped). Lanes share the primes round-robin; multiples
   go to bm, pseudoprime classes to rm (both 64 words of LDS, zeroed by the caller), so
   each prime's offsets are computed once. With `classes`, also returns the lane's word of
   pseudoprime-class members (the candidates this sieve leaves to the strong PRP test).
   Called by all 64 lanes. */
__device__ __forceinline__ u64 deep_sieve(u64 *bm, u64 *rm, u64 group, const DeepPrime *deep,
                                          u32 ndeep, u32 lane, u64 *classes = nullptr)
{
    for (u32 j = lane; j < ndeep; j += 64) {
        const DeepPrime d = deep[j];
        u32 i, ir;
        deep_offsets(group, d.ell, d.m, d.c32_m, d.bar_ell, d.bar_m, &i, &ir);
        if (group <= d.ell && group + 2ull * i == d.ell) i += d.ell; /* never ell itself */
        for (; i < 4096; i += d.ell) atomicOr((unsigned long long *)&bm[i >> 6], 1ull << (i & 63));
        for (; ir < 4096; ir += d.m >> 1)
            atomicOr((unsigned long long *)&rm[ir >> 6], 1ull << (ir & 63));
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const u64 rw = rm[lane];
    const u64 drop = bm[lane] & ~rw;
    if (classes) *classes = rw;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return drop;
}

/* Append one record; `tag` goes to the record's pad field (the range slot in the
   persistent form, 0 in the plain form). */
__device__ __forceinline__ void put_result(ww_result *res, u32 res_cap, u32 *count, u64 p,
                                           int32_t code, int32_t rem, int32_t quot, int32_t tag)
{
    u32 ix = atomicAdd(count, 1u);
    if (ix < res_cap) {
        ww_result r;
        r.p = p;
        r.result = code;
        r.rem = rem;
        r.quot = quot;
        r.pad = tag;
        res[ix] = r;
    }
}

/* Wieferich classification of wieferich.cl:247-293 (unsigned 64-bit comparisons). */
__device__ __forceinline__ void classify_w(u64 p, u64 s0, u64 s1, ww_result *res, u32 cap,
                                           u32 *count, int32_t tag)
{
    const u64 T = 1000; /* SPECIAL_THRESHOLD, WW.h:9 */
    if ((s1 == 0 && s0 == 1) || (s1 == p - 1 && s0 == p - 1)) {
        put_result(res, cap, count, p, 1093, 1, 0, tag);
    } else if (s1 <= T) {
        if (s0 == 1) put_result(res, cap, count, p, 1093, 1, (int32_t)s1, tag);
        else if (s0 == p - 1) put_result(res, cap, count, p, 1093, -1, (int32_t)(s1 + 1), tag);
    } else if (s1 >= p - T) {
        if (s0 == 1) put_result(res, cap, count, p, 1093, 1, (int32_t)(u32)(s1 - p), tag);
        else if (s0 == p - 1)
            put_result(res, cap, count, p, 1093, -1, (int32_t)(u32)(s1 + 1 - p), tag);
    }
}

/* Wall-Sun-Sun classification of wallsunsun.cl:325-349. */
__device__ __forceinline__ void classify_wss(u64 p, u64 c21, ww_result *res, u32 cap,
                                             u32 *count, int32_t tag)
{
    const u64 T = 1000;
    if (c21 == 0) put_result(res, cap, count, p, 1, 0, 0, tag);
    else if (c21 <= T) put_result(res, cap, count, p, 2, 0, (int32_t)c21, tag);
    else if (c21 >= p - T) put_result(res, cap, count, p, 2, 0, (int32_t)(u32)(c21 - p), tag);
}

/* Both tests on one list entry per lane; returns the entry's checksum contribution
   low32(s1) + low32(c21) (wieferich.cl:315, wallsunsun.cl:352), 0 on lanes with act
   false. The ladders run from bounds common to the wave (the largest leading bits of
   their exponents), so every lane runs the same loops; inactive lanes take part in the
   bounds only. Called by all 64 lanes. */
__device__ __forceinline__ u64 test_prime_wave(u64 p, bool act, ww_result *res, u32 cap,
                                               u32 *count, int32_t tag)
{
    const u64 pb = act ? p : 131; /* any prime > 5 for the bounds */
    const int top_w = wave_max(63 - clz64(pb >> 1));
    const int top_f = wave_max(wallsunsun_top(pb));
    /* six table steps when every active lane allows them, else five (wave-uniform) */
    const int skip = __ballot(act && wallsunsun_skip(wallsunsun_n(p), top_f) < 6) ? 5 : 6;
    u64 v = 0;
    if (act) {
        const u64 q = inv64(p);
        const double pinv = 1.0 / (double)p;
        const m2 r2 = m2_r2(m2_one(p), p, pinv);
        const m2 s = wieferich_residue_w(p, q, r2, pinv, top_w);
        const m2 f = wallsunsun_residue_w(p, q, r2, top_f, skip);
        classify_w(p, s.d0, s.d1, res, cap, count, tag);
        classify_wss(p, f.d1, res, cap, count, tag);
        v = (s.d1 & 0xffffffffu) + (f.d1 & 0xffffffffu);
    }
    return v;
}

} // namespace ww

#endif
