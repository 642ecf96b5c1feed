/*
 * ww_kernels.h -- device-side layout shared by the kernels and the C-ABI layer.
 */
#ifndef WW_KERNELS_H
#define WW_KERNELS_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ww_api.h"

namespace ww {

/* Counters in device memory. Field meaning follows the reference's buffers:
   checksum / totalprimecount / resultcount (main.cpp:1446-1475, cleared by
   clearresult.cl:12-26), primecount[0..1] (d_segprimecount, clearn.cl:12-21). */
struct DevState {
    uint64_t checksum;
    uint64_t totalprimecount;
    uint32_t primecount;     /* PRPs appended to the list since the last clear */
    uint32_t primecount_max; /* largest primecount seen by a test launch */
    uint32_t resultcount;    /* result records produced (may exceed capacity) */
    uint32_t reserved;
};

/* Sieve geometry: one 64-bit word of odd numbers (128 integers) per lane, one wave
   per 64 words, four waves per workgroup. */
constexpr int kSieveWaves = 4;
constexpr int kSieveThreads = 64 * kSieveWaves;
constexpr uint64_t kIntsPerWord = 128;
constexpr uint64_t kIntsPerWave = 64 * kIntsPerWord;               /* 8192 */
constexpr uint64_t kIntsPerBlock = kSieveWaves * kIntsPerWave;     /* 32768 */
/* At most 36 of 64 consecutive odd numbers are prime to 15 (4*8 + 4). */
constexpr int kMaxCandPerWord = 36;
constexpr int kWaveCandCap = 64 * kMaxCandPerWord;                 /* 2304 */

constexpr int kTestThreads = 256;

/* One prime of the deep sieve (127 <= ell < 2^16) with its base-2 pseudoprime class
   rho = ell mod m, m = ell * lcm(2, ord_ell(2)) < 2^32, and the reduction constants
   c32_m = 2^32 mod m, bar_ell = floor(2^32 / ell), bar_m = floor(2^64 / m). Same layout
   as ww_deep_prime of the C ABI, which builds the table on the host. */
struct DeepPrime {
    uint32_t ell, m, rho, c32_m, bar_ell, pad;
    uint64_t bar_m;
};
constexpr uint32_t kMaxDeepBound = 65535;

hipError_t launch_sieve_prp(uint64_t lo, uint64_t hi, uint64_t *list, uint64_t cap,
                            DevState *st, const DeepPrime *deep, uint32_t ndeep,
                            const uint64_t *pre, hipStream_t s, const uint64_t *prer = nullptr);
hipError_t launch_ww_tests(const uint64_t *list, uint64_t cap, DevState *st,
                           ww_result *res, uint32_t res_cap, int grid, hipStream_t s);
int tests_blocks_per_cu();
hipError_t launch_residues(const uint64_t *p, uint64_t n, uint64_t *out, hipStream_t s);
hipError_t launch_prp_flags(const uint64_t *n, uint64_t count, uint8_t *out, hipStream_t s);

} // namespace ww

#endif
