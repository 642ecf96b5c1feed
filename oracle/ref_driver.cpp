/*
 * ref_driver.cpp -- drives the reference's own CPU arithmetic, compiled from its
 * sources where they lie (prime.h, wwcpu.cpp, WW.h under the reference tree), into
 * oracle/_ref/libwwref.so. TEST INFRASTRUCTURE ONLY: used to pin the oracle and as
 * the "reference" CPU baseline of bench.py. Nothing of the reference is copied here;
 * this file only calls montInit/strong_prp (prime.h:46-112) and
 * wieferichCPU/wallsunsunCPU (wwcpu.cpp:253-396) in the loop of main.cpp:648-697.
 */
#include <stdint.h>
#include <thread>
#include <vector>

#include "prime.h"
#include "WW.h"

namespace {

struct Sum {
    uint64_t prp = 0, checksum = 0, results = 0;
};

bool no_small_factor(uint64_t n)
{
    static const unsigned small[29] = {3,  5,  7,  11, 13, 17, 19, 23, 29, 31,
                                       37, 41, 43, 47, 53, 59, 61, 67, 71, 73,
                                       79, 83, 89, 97, 101, 103, 107, 109, 113};
    for (unsigned d : small)
        if (n % d == 0) return false;
    return true;
}

Sum run(uint64_t lo, uint64_t hi)
{
    Sum s;
    uint64_t n = lo | 1;
    if (n < lo) return s;
    for (; n < hi; n += 2) {
        if (no_small_factor(n)) {
            uint64_t q, one, r2, nmo;
            montInit(n, &q, &one, &nmo, &r2);
            if (strong_prp(2, n, q, one, nmo, r2)) {
                ++s.prp;
                s.results += (uint64_t)wieferichCPU(n, s.checksum);
                s.results += (uint64_t)wallsunsunCPU(n, s.checksum);
            }
        }
        if (n + 2 < n) break;
    }
    return s;
}

} // namespace

extern "C" {

/* testKernels' CPU loop (main.cpp:648-697) over [lo, hi) on `threads` threads. */
void ref_range(uint64_t lo, uint64_t hi, int threads, uint64_t out[3])
{
    if (threads < 1) threads = 1;
    std::vector<Sum> part((size_t)threads);
    std::vector<std::thread> pool;
    uint64_t step = (hi - lo) / (uint64_t)threads;
    for (int i = 0; i < threads; ++i) {
        uint64_t a = lo + step * (uint64_t)i;
        uint64_t b = (i == threads - 1) ? hi : lo + step * (uint64_t)(i + 1);
        pool.emplace_back([&part, i, a, b] { part[(size_t)i] = run(a, b); });
    }
    Sum t;
    for (int i = 0; i < threads; ++i) {
        pool[(size_t)i].join();
        t.prp += part[(size_t)i].prp;
        t.checksum += part[(size_t)i].checksum;
        t.results += part[(size_t)i].results;
    }
    out[0] = t.prp;
    out[1] = t.checksum;
    out[2] = t.results;
}

/* Base-2 strong PRP of the reference (prime.h:46-112). */
int ref_strong_prp2(uint64_t n)
{
    uint64_t q, one, r2, nmo;
    montInit(n, &q, &one, &nmo, &r2);
    return strong_prp(2, n, q, one, nmo, r2);
}

/* Checksum slice and flag of each test for one p (wwcpu.cpp:253-396). */
int ref_wieferich(uint64_t p, uint64_t *slice)
{
    uint64_t c = 0;
    int f = wieferichCPU(p, c);
    *slice = c;
    return f;
}

int ref_wallsunsun(uint64_t p, uint64_t *slice)
{
    uint64_t c = 0;
    int f = wallsunsunCPU(p, c);
    *slice = c;
    return f;
}

} // extern "C"
