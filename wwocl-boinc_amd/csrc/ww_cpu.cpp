/*
 * ww_cpu.cpp -- host-side checks of the drop-in (CPU): the validation the reference
 * runs on the host around the device path, with the same arithmetic as the kernels
 * (ww_arith.h compiled for the host).
 *   ww_cpu_is_prime   isPrime, main.cpp:301-324
 *   ww_cpu_is_prp     isPRP, main.cpp:245-291
 *   ww_cpu_range      testKernels' CPU loop, main.cpp:648-697
 */
#include <vector>

#include "../../include/ww_api.h"
#include "ww_arith.h"

using namespace ww;

namespace {

/* Strong probable-prime test of odd n > 1 to base `base` (prime.h:66-112). */
bool strong_prp(u64 base, u64 n)
{
    u64 q = inv64(n);
    u64 one = mont_one(n);
    u64 mone = n - one;
    u64 two = add_mod(one, one, n);
    u64 r2 = add_mod(two, two, n); /* 4 */
    for (int i = 0; i < 5; ++i) r2 = mont_mul(r2, r2, n, q); /* 4^(2^5) = 2^64 */
    u64 nm1 = n - 1;
    int t = ctz64(nm1);
    u64 d = nm1 >> t;
    u64 mb = mont_mul(base, r2, n, q);
    u64 a = mb;
    for (int i = 62 - clz64(d); i >= 0; --i) {
        a = mont_mul(a, a, n, q);
        if ((d >> i) & 1) a = mont_mul(a, mb, n, q);
    }
    if (a == one || a == mone) return true;
    for (int s = 1; s < t; ++s) {
        a = mont_mul(a, a, n, q);
        if (a == mone) return true;
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

} // namespace

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

int ww_cpu_range(uint64_t low, uint64_t high, uint64_t out[3])
{
    if (!out) return WW_E_ARG;
    const u64 T = 1000; /* SPECIAL_THRESHOLD, WW.h:9 */
    u64 prp = 0, sum = 0, flags = 0;
    u64 n = low | 1;
    if (n >= low) {
        for (; n < high; n += 2) {
            if (no_factor_to_113(n) && spsp2(n, inv64(n), mont_one(n))) {
                ++prp;
                u64 q = inv64(n);
                m2 one = m2_one(n);
                m2 s = wieferich_residue(n, q, one);
                m2 f = wallsunsun_residue(n, q, one);
                sum += (s.d1 & 0xffffffffu) + (f.d1 & 0xffffffffu);
                /* wieferichCPU / wallsunsunCPU return values (wwcpu.cpp:290-303, 382-395) */
                if ((s.d1 == 0 && s.d0 == 1) || (s.d1 == n - 1 && s.d0 == n - 1) || s.d1 <= T ||
                    s.d1 >= n - T)
                    ++flags;
                if (f.d1 <= T || f.d1 >= n - T) ++flags;
            }
            if (n + 2 < n) break;
        }
    }
    out[0] = prp;
    out[1] = sum;
    out[2] = flags;
    return WW_OK;
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

} // extern "C"
