/*
 * ww_cpu.cpp -- host-side checks of the drop-in (CPU): the validation the reference
 * runs on the host around the device path, with the same arithmetic as the kernels
 * (ww_arith.h compiled for the host).
 *   ww_cpu_is_prime   isPrime, main.cpp:301-324
 *   ww_cpu_is_prp     isPRP, main.cpp:245-291
 *   ww_cpu_range      testKernels' CPU loop, main.cpp:648-697
 */
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

} // extern "C"
