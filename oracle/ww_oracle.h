/*
 * ww_oracle.h -- CPU oracle for the Wieferich / Wall-Sun-Sun search path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/, the smoke()
 * entry and the cpu_baseline leg of bench.py may load it. The product path
 * (libww.so and the wwocl host) never links or calls it.
 *
 * It restates, in plain C, the reference CPU path of WWocl:
 *   prime.h:7-112        64-bit Montgomery arithmetic and the base-2 strong PRP test
 *   wwcpu.cpp:19-396     mod-p^2 "double precision" Montgomery arithmetic, the
 *                        Wieferich and Wall-Sun-Sun tests and their checksum slices
 *   main.cpp:245-324     isPRP / isPrime used to validate results
 *   main.cpp:528-587     findWheelOffset
 *   main.cpp:648-697     the CPU range loop (odd N, trial division 3..113, spsp(2), W, WSS)
 *   kernels/wieferich.cl:247-293, kernels/wallsunsun.cl:325-349
 *                        the device-side result classification (result code, rem, quot)
 * plus a textbook segmented sieve of Eratosthenes standing in for primesieve
 * (main.cpp:498 uses primesieve_generate_primes as the prime list to compare against).
 */
#ifndef WW_ORACLE_H
#define WW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One result record; same layout as ww_result in include/ww_api.h. */
typedef struct {
    uint64_t p;
    int32_t result; /* 1093 = Wieferich, 1 = WSS prime, 2 = WSS special instance */
    int32_t rem;
    int32_t quot;
    int32_t pad;
} or_result;

/* Summary of one range, mirroring the counters of main.cpp:602-605. */
typedef struct {
    uint64_t prp_count;   /* list entries: spsp(2) with no prime factor <= 113 */
    uint64_t checksum;    /* sum of low32(s1) + low32(c21), mod 2^64 */
    uint64_t result_count;/* device-style result records (wieferich.cl / wallsunsun.cl) */
    uint64_t cpu_result_count; /* wieferichCPU()+wallsunsunCPU() return values (main.cpp:689-694) */
} or_summary;

uint64_t or_invert(uint64_t p);
uint64_t or_mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q);
int or_strong_prp(uint64_t base, uint64_t n);
int or_is_candidate(uint64_t n);
int or_is_prp(uint64_t n);
int or_is_prime(uint64_t n);

void or_m2p_one(uint64_t p, uint64_t out[2]);
void or_wieferich(uint64_t p, uint64_t out[2]);   /* s0, s1 */
void or_wallsunsun(uint64_t p, uint64_t out[2]);  /* c20, c21 */
int or_wieferich_cpu_flag(uint64_t p, uint64_t *checksum);  /* wwcpu.cpp:253-304 */
int or_wallsunsun_cpu_flag(uint64_t p, uint64_t *checksum); /* wwcpu.cpp:306-396 */
int or_classify_wieferich(uint64_t p, uint64_t s0, uint64_t s1, or_result *rec);
int or_classify_wallsunsun(uint64_t p, uint64_t c21, or_result *rec);

void or_find_wheel_offset(uint64_t *start, int32_t *index);

/* The reference CPU loop over [lo, hi); results (sorted by p, Wieferich first) go to
   out[0..cap); returns the number of records written (<= cap). */
uint64_t or_range(uint64_t lo, uint64_t hi, or_summary *sum, or_result *out, uint64_t cap);
/* Same loop split over `threads` std-thread-free pthreads; results are not collected. */
void or_range_mt(uint64_t lo, uint64_t hi, int threads, or_summary *sum);

/* List of candidates (odd, no factor 3..113) passing spsp(2), ascending. */
uint64_t or_prp_list(uint64_t lo, uint64_t hi, uint64_t *out, uint64_t cap);
/* All primes in [lo, hi), ascending, by a segmented sieve of Eratosthenes. */
uint64_t or_sieve_primes(uint64_t lo, uint64_t hi, uint64_t *out, uint64_t cap);

#ifdef __cplusplus
}
#endif
#endif
