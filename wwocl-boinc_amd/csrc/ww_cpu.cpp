/*
 * ww_cpu.cpp -- the device arithmetic built for the host, for verification, and the
 * host-side sieve tables:
 *   ww_cpu_residues, ww_cpu_spsp2   ww_arith.h's ladders and windowed PRP test on the CPU,
 *                                   with the bounds a wave would share as arguments
 *   ww_deep_table, ww_presieve_table  the deep-sieve and pre-sieve primes with their
 *                                   base-2 pseudoprime classes
 *   ww_cpu_bucket_hits, ww_cpu_order2  ww_bucket.h's walk of one bucket-sieve prime and
 *                                   its base-2 order
 * The drop-in's validation and self-test (isPrime, isPRP, the CPU range loop) are in
 * ww_selftest.cpp, in a formulation independent of this arithmetic.
 */
#include <vector>

#include "../../include/ww_api.h"
#include "ww_arith.h"
#include "ww_bucket.h"

using namespace ww;

static uint32_t powmod32(uint32_t b, uint32_t e, uint32_t m)
{
    uint64_t r = 1 % m, x = b % m;
    for (; e; e >>= 1) {
        if (e & 1) r = r * x % m;
        x = x * x % m;
    }
    return (uint32_t)r;
}

int ww_set_error(int code, const char *fmt, ...); /* ww_api.cpp */

namespace ww {

const SmallDiv *small_div_table()
{
    static const std::vector<SmallDiv> tab = [] {
        std::vector<SmallDiv> t;
        std::vector<uint8_t> comp(1u << 16, 0);
        for (uint32_t q = 3; q < (1u << 16); q += 2) {
            if (comp[q]) continue;
            for (uint32_t j = q * q; j < (1u << 16); j += 2 * q) comp[j] = 1;
            uint32_t inv = q; /* q q = 1 (mod 8); each Newton step doubles the good bits */
            for (int k = 0; k < 4; ++k) inv *= 2u - q * inv;
            t.push_back({q, inv, 0xffffffffu / q, 0});
        }
        return t;
    }();
    return tab.data();
}

} // namespace ww

extern "C" {

int ww_cpu_residues(const uint64_t *p, uint64_t n, int32_t top_w, int32_t top_f,
                    int32_t wss_skip, uint64_t *out)
{
    if (n && (!p || !out)) return ww_set_error(WW_E_ARG, "null buffer");
    for (uint64_t i = 0; i < n; ++i) {
        const u64 pi = p[i];
        if (pi < 127 || (pi & 1) == 0 || mod5(pi) == 0)
            return ww_set_error(WW_E_ARG, "p[%llu] = %llu is not odd, >= 127 and prime to 5",
                                (unsigned long long)i, (unsigned long long)pi);
        const int lw = 63 - clz64(pi >> 1), lf = wallsunsun_top(pi);
        const int tw = top_w < 0 ? lw : top_w, tf = top_f < 0 ? lf : top_f;
        if (tw < lw || tw > 63 || tf < lf || tf > 63)
            return ww_set_error(WW_E_ARG, "ladder bounds %d, %d outside [%d, 63], [%d, 63]", tw, tf,
                                lw, lf);
        const int most = wallsunsun_skip(wallsunsun_n(pi), tf);
        const int skip = wss_skip < 0 ? most : wss_skip;
        if (skip != 5 && skip != most)
            return ww_set_error(WW_E_ARG, "wss_skip %d: this p and bound allow 5%s", skip,
                                most == 6 ? " or 6" : "");
        const u64 q = inv64(pi);
        const double pinv = 1.0 / (double)pi;
        const m2 r2 = m2_r2(m2_one(pi), pi, pinv);
        const m2 s = wieferich_residue_w(pi, q, r2, pinv, tw);
        const m2 f = wallsunsun_residue_w(pi, q, r2, tf, skip);
        out[4 * i] = s.d0;
        out[4 * i + 1] = s.d1;
        out[4 * i + 2] = f.d0;
        out[4 * i + 3] = f.d1;
    }
    return WW_OK;
}

int ww_cpu_spsp2(uint64_t n, int32_t top, int32_t tmax)
{
    if (n < 3 || (n & 1) == 0) return ww_set_error(WW_E_ARG, "n = %llu is not odd and > 1",
                                                   (unsigned long long)n);
    const u64 e = n - 1;
    const int lead = 63 - clz64(e), t = ctz64(e);
    if (top < 0) top = lead;
    if (tmax < 0) tmax = t;
    if (top < lead || top > 63 || tmax < t || tmax > top)
        return ww_set_error(WW_E_ARG, "bounds top %d, tmax %d outside [%d, 63], [%d, top]", top,
                            tmax, lead, t);
    return spsp2_window(n, inv64(n), mont_one(n), top, tmax) ? 1 : 0;
}

uint32_t ww_deep_table(uint32_t bound, ww_deep_prime *out, uint32_t cap)
{
    if (bound > 65535) bound = 65535; /* m = ell * lcm(2, ord) must stay below 2^32 */
    std::vector<uint8_t> comp(bound + 1, 0);
    uint32_t n = 0;
    for (uint32_t i = 2; i <= bound; ++i) {
        if (comp[i]) continue;
        for (uint64_t j = (uint64_t)i * i; j <= bound; j += i) comp[j] = 1;
        if (i <= 113) continue;
        if (n < cap && out) {
            /* ord_ell(2) divides ell - 1: strip each prime factor while 2^(ord/f) = 1 */
            uint32_t ord = i - 1, rest = i - 1;
            for (uint32_t f = 2; rest > 1; ++f) {
                if (f * f > rest) f = rest;
                if (rest % f) continue;
                while (rest % f == 0) rest /= f;
                while (ord % f == 0 && powmod32(2, ord / f, i) == 1) ord /= f;
            }
            const uint64_t L = (ord % 2 == 0) ? ord : 2 * (uint64_t)ord;
            ww_deep_prime &d = out[n];
            d.ell = i;
            d.m = (uint32_t)(i * L);
            /* rho = 0 mod ell and 1 mod L; ell - 1 is a multiple of ord and even, so
               ell = 1 mod L and rho = ell */
            d.rho = i;
            d.c32_m = (uint32_t)((1ull << 32) % d.m);
            d.bar_ell = (uint32_t)(0xffffffffu / i);  /* floor(2^32 / ell), ell odd */
            d.pad = 0;
            d.bar_m = ~0ull / d.m;                     /* floor(2^64 / m), m no power of 2 */
        }
        ++n;
    }
    return n < cap ? n : cap;
}

uint32_t ww_presieve_table(uint32_t above, uint32_t bound, ww_wide_prime *out, uint32_t cap)
{
    if (bound > WW_MAX_PRESIEVE_BOUND) bound = WW_MAX_PRESIEVE_BOUND;
    if (above < 113) above = 113;
    if (bound <= above) return 0;
    /* smallest prime factor of every k <= bound, so that ell - 1 factors at once */
    std::vector<uint32_t> spf(bound + 1, 0);
    for (uint32_t i = 2; i <= bound; ++i)
        if (!spf[i])
            for (uint64_t j = i; j <= bound; j += i)
                if (!spf[j]) spf[j] = i;
    uint32_t n = 0;
    for (uint32_t i = above + 1; i <= bound && n < cap; ++i) {
        if (spf[i] != i) continue;
        /* ord_ell(2) divides ell - 1: strip each prime factor while 2^(ord/f) = 1 */
        uint32_t ord = i - 1, rest = i - 1;
        while (rest > 1) {
            const uint32_t f = spf[rest];
            while (rest % f == 0) rest /= f;
            while (ord % f == 0 && powmod32(2, ord / f, i) == 1) ord /= f;
        }
        const uint64_t L = (ord % 2 == 0) ? ord : 2 * (uint64_t)ord;
        ww_wide_prime &d = out[n++];
        d.ell = i;
        d.pad = 0;
        d.m = i * L;
        d.bar_ell = ~0ull / i; /* floor(2^64 / ell), ell odd */
        d.bar_m = ~0ull / d.m; /* floor(2^64 / m), m no power of 2 */
    }
    return n;
}

int ww_cpu_bucket_hits(uint64_t g0, uint64_t nbits, uint32_t ell, uint64_t *bits,
                       uint8_t *rescue, uint64_t cap, uint64_t *count)
{
    if (!count || (cap && (!bits || !rescue))) return ww_set_error(WW_E_ARG, "null buffer");
    if ((g0 & 1) == 0 || ell <= (1u << 20) || (ell & 1) == 0)
        return ww_set_error(WW_E_ARG, "g0 must be odd and ell an odd prime above 2^20");
    /* the window's cofactors k (l k in [g0, g0 + 2 nbits)) number at most 2 nbits / l + 1 */
    if (nbits > (kWalkMaxK - 1) * ell / 2)
        return ww_set_error(WW_E_ARG, "window of %llu odd numbers spans more than %llu multiples",
                            (unsigned long long)nbits, (unsigned long long)kWalkMaxK);
    static constexpr WheelTables wt = make_wheel();
    const u32 ord = order2_u32(ell, small_div_table(), kSmallOddPrimes);
    BucketWalk w = bucket_start(g0, ell, inv_below(ell), ord, wt.res, wt.bits, wt.rank);
    uint64_t n = 0;
    for (; w.d < 2 * nbits; bucket_step(w, ell, ord, wt.gap)) {
        if (!bucket_visit(w, wt.visit_a, wt.visit_b, wt.visit_c)) continue;
        if (n < cap) {
            bits[n] = w.d >> 1;
            rescue[n] = w.c == 0u;
        }
        ++n;
    }
    *count = n;
    return WW_OK;
}

int ww_cpu_order2(uint32_t ell, uint32_t *order)
{
    if (!order) return ww_set_error(WW_E_ARG, "null output");
    if (ell <= (1u << 20) || (ell & 1) == 0)
        return ww_set_error(WW_E_ARG, "ell must be an odd prime above 2^20");
    *order = order2_u32(ell, small_div_table(), kSmallOddPrimes);
    return WW_OK;
}

} // extern "C"
