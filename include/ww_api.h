/*
 * ww_api.h -- C ABI of libww, the MI355X (gfx950) search path for Wieferich and
 * Wall-Sun-Sun primes.
 *
 * This is the drop-in boundary that replaces the reference's device layer: the
 * simpleCL wrapper (simpleCL.h:63-133), the raw OpenCL buffer/queue calls of main.cpp
 * and the five kernel entry points with their argument contracts (main.cpp:1596-1629).
 * Plain C types only; every entry point returns 0 on success or a negative WW_E* code,
 * with a message available from ww_last_error() (the reference prints and exits on
 * every device error, simpleCL.c:30-348; the drop-in host maps a nonzero return to
 * that behaviour).
 *
 * Semantics kept from the reference:
 *   - ranges are half-open [low, high) over unsigned 64-bit integers;
 *   - the PRP list holds every n in the range that is odd, has no prime factor in
 *     3..113 and is a base-2 strong probable prime (getsegprimes.cl:228-301);
 *   - each list entry p adds low32(s1) of s = 2^((p-1)/2) mod p^2 = s0 + p*s1 and
 *     low32(c21) of F_{p-(p/5)} mod p^2 = c20 + p*c21 to a 64-bit checksum
 *     (wieferich.cl:315, wallsunsun.cl:352);
 *   - result records carry result code 1093 (Wieferich), 1 (Wall-Sun-Sun prime) or
 *     2 (Wall-Sun-Sun special instance) with the rem/quot conventions of
 *     wieferich.cl:247-293 and wallsunsun.cl:325-349.
 * Changed on purpose: the list stores p only (p^-1 mod 2^64 and 2^64 mod p^2 are
 * recomputed in registers), and the Wieferich and Wall-Sun-Sun kernels are one
 * launch, so there is no g_one round trip.
 *
 * Thread safety: calls on different contexts are independent. Every entry point that
 * takes a context holds that context's lock for its duration, so calls on one context
 * from several host threads are serialised (they still share one in-order stream, as the
 * reference's single queue does, main.cpp:1683-1795). A batch is used under its
 * context's lock; ww_batch_set_grid and ww_batch_destroy take no context and must not
 * race with a launch of the same batch. Servers lock themselves.
 */
#ifndef WW_API_H
#define WW_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WW_OK 0
#define WW_E_ARG (-1)      /* invalid argument */
#define WW_E_HIP (-2)      /* HIP runtime error */
#define WW_E_NODEV (-3)    /* no usable GPU */
#define WW_E_CAPACITY (-4) /* request exceeds a buffer capacity */

/* Smallest START the search accepts (main.cpp:1265-1269 clamps to it). */
#define WW_MIN_START 127ull
/* Result-buffer size of the reference (numresults, main.cpp:1119). */
#define WW_DEFAULT_RESULTS 15000u
/* Default bound of the deep sieve (see ww_set_deep_bound). */
#define WW_DEFAULT_DEEP_BOUND 2048u
/* Default bound of the wide pre-sieve (see ww_set_presieve_bound). */
#define WW_DEFAULT_PRESIEVE_BOUND 1048576u

/* One result record (the four arrays gpuSpecialPrime/gpuResult/gpuRem/gpuQuot of
   main.cpp:1458-1493, as one struct). */
typedef struct {
    uint64_t p;
    int32_t result; /* 1093 Wieferich, 1 WSS prime, 2 WSS special instance */
    int32_t rem;
    int32_t quot;
    int32_t pad;
} ww_result;

/* Device counters read back at a checkpoint (main.cpp:793-796). */
typedef struct {
    uint64_t checksum;        /* gpuChecksum */
    uint64_t totalprimecount; /* d_totalprimecount: PRPs tested since the last clear */
    uint32_t primecount;      /* d_segprimecount[0]: current list length */
    uint32_t primecount_max;  /* d_segprimecount[1]: largest list length tested */
    uint32_t resultcount;     /* gpuResultCount (may exceed the result capacity) */
    uint32_t reserved;
} ww_counters;

typedef struct {
    char name[256];
    char arch[64];
    int compute_units;
    int clock_khz;
    uint64_t global_mem;
    int warp_size;
} ww_device_info_t;

typedef struct ww_ctx ww_ctx;

/* Message of the last failing call on this thread. */
const char *ww_last_error(void);

/* Number of visible GPUs (0 when none). */
int ww_device_count(int *count);

/* Device properties (replaces the clGetDeviceInfo queries, main.cpp:1325-1423). */
int ww_device_info(int device, ww_device_info_t *info);

/* Open a context on `device` with a PRP list of `list_capacity` entries and
   `result_capacity` result records. Replaces clCreateContext / queue creation and the
   buffer allocations of main.cpp:1310-1323, 1446-1493, 1563-1581. Counters start
   cleared. */
int ww_open(int device, uint64_t list_capacity, uint32_t result_capacity, ww_ctx **out);

/* Release the context and its device memory (cleanup, main.cpp:69-90). Destroy its
   batches and servers first. */
int ww_close(ww_ctx *ctx);

/* List capacity for one range [low, high): the Montgomery-Vaughan (Brun-Titchmarsh)
   bound 2w/ln(w) on the primes in a window of width w, which is a proven bound, times
   1.02 plus 4096 for the base-2 strong pseudoprimes the window may hold. That slack is a
   heuristic, not a bound (there are about 3.1e7 such pseudoprimes below 2^64, and no
   window is known to hold more than a few): a list that overflows anyway is caught by the
   counters (primecount > capacity, main.cpp:802-805, as the drop-in host checks after
   every chunk). Replaces the primesieve-based sizing (main.cpp:1087-1091). */
uint64_t ww_list_bound(uint64_t low, uint64_t high);

/* Bound B (0 or 127..65535) of the deep sieve applied by the candidate generators after
   the reference's own filter by 3..113. No counterpart in the reference: it removes
   composites with a factor in 127..B *except* the single residue class per factor in
   which base-2 pseudoprimes divisible by it must lie, so the PRP list is unchanged and
   only fewer composites reach the strong PRP test. B = 0 turns it off. Takes effect for
   launches issued afterwards. Default WW_DEFAULT_DEEP_BOUND. */
int ww_set_deep_bound(ww_ctx *ctx, uint32_t bound);
int ww_get_deep_bound(ww_ctx *ctx, uint32_t *bound);

/* Bound B2 (0 = off, else at most WW_MAX_PRESIEVE_BOUND) of the wide pre-sieve of the
   persistent form: before each batch launch, one kernel drops, in chunks of 2^21 integers,
   the multiples of the primes in (deep bound, B2] except their pseudoprime classes (the
   same exact rule as the deep sieve), so fewer composites reach the strong PRP test.
   Applies to batches created afterwards (each holds 512 bytes per 8192-integer segment)
   and to the plain form's chunks (ww_getsegprimes keeps one such buffer per context).
   No counterpart in the reference. Default WW_DEFAULT_PRESIEVE_BOUND; a larger bound
   than before rebuilds the prime table on the host. */
int ww_set_presieve_bound(ww_ctx *ctx, uint32_t bound);
int ww_get_presieve_bound(ww_ctx *ctx, uint32_t *bound);

/* Smallest range width that gets the bucket sieve in batches created afterwards (see
   ww_set_bucket_sieve). */
#define WW_DEFAULT_BUCKET_MIN_WIDTH (1ull << 30)

/* The bucket sieve of the persistent form, for ranges at least min_width wide in
   batches created afterwards and for chunks of ww_search_fused and ww_getsegprimes at
   least that wide (UINT64_MAX: none). No counterpart in the reference. Such a
   range is sieved by every prime up to sqrt(hi - 1): the deep sieve and the wide pre-sieve
   take the primes up to the pre-sieve bound, and one more kernel drops the hits of the
   primes above it, up to 2^32, into per-chunk buckets that the pre-sieve applies. Every
   tier spares (and flags) the base-2 pseudoprime class of each prime, so a candidate that no
   tier dropped or flagged is a prime and skips the strong PRP test; a flagged one is tested
   as before. The PRP list, checksums and records are unchanged. It takes effect where the
   deep bound is at least 1024, the pre-sieve bound at least 2^20 and the pre-sieve segment
   2^20 bits (the defaults); elsewhere such a batch runs as without it. Costs per range
   about 0.67 bytes of device memory per integer of width (5.7 GB for a 2^33-wide range;
   single chunks keep one such store per context, for the widest chunk so far) and, on
   first use, the odd primes below 2^32 and their base-2 orders, 813 MB each, built on the
   device in about 0.05 s. Default WW_DEFAULT_BUCKET_MIN_WIDTH. */
int ww_set_bucket_sieve(ww_ctx *ctx, uint64_t min_width);
int ww_get_bucket_sieve(ww_ctx *ctx, uint64_t *min_width);

/* Scale (0 < scale <= 16, default 1) of the bucket capacities of batches created
   afterwards. Below 1 buckets overflow: the chunks concerned then test all their
   candidates, which is exact but slower, so this is for testing that path. */
int ww_set_bucket_capacity(ww_ctx *ctx, double scale);

/* The context's table of the odd primes below 2^32 (built on first use, 203,280,220
   entries): copies entries [first, first + n) to out (host) and writes the entry count to
   *total. Verification; blocking. */
int ww_prime_table(ww_ctx *ctx, uint64_t first, uint64_t n, uint32_t *out, uint64_t *total);

/* The same entries of the context's order table: ord_l(2) of each table prime l above 2^20
   (the class test of the bucket sieve), 0 for the smaller ones. Verification; blocking. */
int ww_prime_orders(ww_ctx *ctx, uint64_t first, uint64_t n, uint32_t *out, uint64_t *total);

/* The bucket sieve's walk of one sieving prime, on the host with the code the device runs
   (verification): the multiples n = ell*k in [g0, g0 + 2 nbits) with k prime to
   30030 x 17 x 19 x ... x 53 (g0 odd, 2^20 < ell < 2^32 prime, and 2 nbits <= 30028 ell: the
   window holds fewer than 30030 cofactors, as every bucket span's does), as bit indices
   (n - g0) / 2 with a 1 in rescue[] where 2^(n-1) = 1 (mod ell) (the prime's base-2
   pseudoprime class). Writes up to cap of each; *count = the number of hits. */
int ww_cpu_bucket_hits(uint64_t g0, uint64_t nbits, uint32_t ell, uint64_t *bits,
                       uint8_t *rescue, uint64_t cap, uint64_t *count);

/* ord_ell(2) for an odd prime 2^20 < ell < 2^32, on the host with the code the order table's
   kernel runs (verification). */
int ww_cpu_order2(uint32_t ell, uint32_t *order);

/* Launch shape of the persistent kernel for launches issued afterwards: waves per
   workgroup (2, 4 or 8) and the wide pre-sieve's segment, the LDS bitmap one workgroup
   marks at a time, as a bit count (15..20: 2^15 to 2^20 odd numbers). No counterpart in
   the reference; results do not depend on either. Defaults 4 and 20. */
int ww_set_launch_shape(ww_ctx *ctx, uint32_t waves_per_group, uint32_t segment_bits);
int ww_get_launch_shape(ww_ctx *ctx, uint32_t *waves_per_group, uint32_t *segment_bits);

/* clearresult (clearresult.cl:12-26): zero checksum, resultcount, totalprimecount
   and primecount_max. Stream-ordered. */
int ww_clear_results(ww_ctx *ctx);

/* clearn (clearn.cl:12-21): empty the PRP list. Stream-ordered. */
int ww_clear_prp_count(ww_ctx *ctx);

/* getsegprimes (getsegprimes.cl:228-301, launched at main.cpp:1719-1727): append the
   PRPs of [low, high) to the list. The reference also passes a mod-30 wheel index
   computed by findWheelOffset (main.cpp:528-587); this sieve needs none. A range at
   least the bucket width (ww_set_bucket_sieve) is sieved up to sqrt(high - 1) first,
   and its candidates that no sieve tier flags join the list without the PRP test.
   Requires WW_MIN_START <= low < high and high - low <= 2^34. Stream-ordered. */
int ww_getsegprimes(ww_ctx *ctx, uint64_t low, uint64_t high);

/* wieferich + wallsunsun (wieferich.cl:196-328, wallsunsun.cl:221-354, launched at
   main.cpp:1738-1741): test every list entry, add to the checksum, append result
   records, add the list length to totalprimecount and update primecount_max.
   Stream-ordered. */
int ww_wieferich_wallsunsun(ww_ctx *ctx);

/* One whole range as the reference's chunk loop does it (main.cpp:1683-1741):
   for each chunk of at most `chunk` integers: clearn, getsegprimes,
   wieferich+wallsunsun. chunk = 0 picks the context's default. Stream-ordered; the
   host does not wait between chunks. */
int ww_search_range(ww_ctx *ctx, uint64_t low, uint64_t high, uint64_t chunk);

/* One chunk [low, high) (width at most 2^34) through the persistent kernel instead of
   getsegprimes + wieferich/wallsunsun: the same checksum, tested count and records are
   added to the context's counters and results (what ww_read_counters and
   ww_read_results return), but no PRP list is written (its count stays as it is, so
   ww_read_prps and the tests kernel do not apply to such a chunk). Stream-ordered; its
   time counts as test-kernel time in ww_kernel_times. A chunk at least the bucket width
   (ww_set_bucket_sieve) is sieved up to sqrt(high - 1) first, and its candidates that no
   sieve tier flags skip the strong PRP test. No counterpart in the reference: a faster
   body for its chunk loop (main.cpp:1683-1741) when the list is not needed. */
int ww_search_fused(ww_ctx *ctx, uint64_t low, uint64_t high);

/* Wait for all queued work (sclFinish / waitOnEvent, simpleCL.c, main.cpp:329-377). */
int ww_finish(ww_ctx *ctx);

/* Blocking read of the counters (main.cpp:793-796). */
int ww_read_counters(ww_ctx *ctx, ww_counters *out);

/* Blocking read of up to `cap` result records (main.cpp:837-840); *n receives the
   number copied. */
int ww_read_results(ww_ctx *ctx, ww_result *out, uint32_t cap, uint32_t *n);

/* Blocking read of the PRP list (checkPRPS, main.cpp:460-470); *n receives the
   list length (entries beyond cap are not copied). */
int ww_read_prps(ww_ctx *ctx, uint64_t *out, uint64_t cap, uint64_t *n);

/* Kernel milliseconds accumulated since the last call (and reset): the sieve/PRP
   kernel and the Wieferich/Wall-Sun-Sun kernel, from HIP events on the context's
   stream (ProfilesclEnqueueKernel, simpleCL.c:511-534). Waits for queued work. */
int ww_kernel_times(ww_ctx *ctx, double *ms_sieve, double *ms_tests, uint64_t *launches);

/* Per-value access to the device arithmetic, for verification. ww_residues writes
   out[4*i..4*i+3] = (s0, s1, c20, c21) for p[i] (each p odd, >= 127, prime to 5);
   ww_prp_flags writes 1 where n[i] (odd, > 1) is a base-2 strong probable prime.
   Host buffers in and out; blocking. */
int ww_residues(ww_ctx *ctx, const uint64_t *p, uint64_t n, uint64_t *out);
int ww_prp_flags(ww_ctx *ctx, const uint64_t *n, uint64_t count, uint8_t *out);

/* Host-side checks the drop-in uses to validate results (CPU, no device needed):
   isPrime (main.cpp:301-324) and isPRP (main.cpp:245-291). */
int ww_cpu_is_prime(uint64_t n);
int ww_cpu_is_prp(uint64_t n);

/* The device arithmetic built for the host, for verification (no counterpart in the
   reference). ww_cpu_residues writes out[4*i..4*i+3] as ww_residues does, running the
   ladders from the given leading-bit bounds (any value from p's own leading bit to 63
   gives the same result, which is what lets a wave share one bound; -1 = p's own) and
   with wss_skip leading Wall-Sun-Sun steps read from the table of V_k(3, 1): 5, or 6 when
   the six leading bits of the ladder's exponent below top_f give k <= 45 (a wave takes 6
   when all its lanes allow it), -1 = the largest this p allows.
   ww_cpu_spsp2 is the windowed strong PRP test of n with bounds top (>= the leading bit
   of n - 1, <= 63) and tmax (>= the trailing zero count of n - 1, <= top), -1 = n's own:
   1 or 0, or WW_E_ARG. */
int ww_cpu_residues(const uint64_t *p, uint64_t n, int32_t top_w, int32_t top_f,
                    int32_t wss_skip, uint64_t *out);
int ww_cpu_spsp2(uint64_t n, int32_t top, int32_t tmax);

/* One deep-sieve prime: ell < 2^16, its base-2 pseudoprime class rho mod m with
   m = ell * lcm(2, ord_ell(2)) < 2^32 (rho = 0 mod ell, rho = 1 mod lcm(2, ord_ell(2)),
   which makes rho = ell), and reduction constants c32_m = 2^32 mod m,
   bar_ell = floor(2^32 / ell), bar_m = floor(2^64 / m). */
typedef struct {
    uint32_t ell, m, rho, c32_m, bar_ell, pad;
    uint64_t bar_m;
} ww_deep_prime;

/* Host-side (CPU) table of the deep-sieve primes 127 <= ell <= bound (bound < 2^16),
   ascending; writes up to cap entries and returns the number written. No counterpart in
   the reference. */
uint32_t ww_deep_table(uint32_t bound, ww_deep_prime *out, uint32_t cap);

/* One prime of the wide pre-sieve (deep primes beyond the per-segment bound, up to
   WW_MAX_PRESIEVE_BOUND): ell, m = ell * lcm(2, ord_ell(2)) < 2^45 (its base-2
   pseudoprime class is ell mod m), and Barrett constants bar_ell = floor(2^64 / ell),
   bar_m = floor(2^64 / m). No counterpart in the reference. */
typedef struct {
    uint32_t ell, pad;
    uint64_t m, bar_ell, bar_m;
} ww_wide_prime;

#define WW_MAX_PRESIEVE_BOUND 4194304u

/* Host-side (CPU) table of the primes above < ell <= bound (bound <= WW_MAX_PRESIEVE_BOUND,
   above >= 113), ascending; writes up to cap entries and returns the number written. */
uint32_t ww_presieve_table(uint32_t above, uint32_t bound, ww_wide_prime *out, uint32_t cap);

/* The reference CPU path over [low, high) (main.cpp:648-697), on the host with the
   same arithmetic: the startup self-test's comparison target (testKernels). out =
   {prp count, checksum, cpu result count}. */
int ww_cpu_range(uint64_t low, uint64_t high, uint64_t out[3]);

/* ---------------------------------------------------------------------------
 * Persistent form: many ranges in one launch. No counterpart in the reference, which
 * runs one range per launch sequence (main.cpp:1683-1741); per-range checksums and
 * counts equal what the plain entry points give for each range alone.
 * ------------------------------------------------------------------------- */

typedef struct {
    uint64_t lo, hi; /* half-open, WW_MIN_START <= lo < hi */
} ww_range;

typedef struct {
    uint64_t checksum;  /* as gpuChecksum for this range alone */
    uint64_t prp_count; /* tested primes (list entries) of this range */
} ww_range_sum;

typedef struct ww_batch ww_batch;

/* Upload a table of n >= 1 ranges and allocate their sums and `result_capacity`
   records. Allocates device memory. */
int ww_batch_create(ww_ctx *ctx, const ww_range *ranges, uint32_t n, uint32_t result_capacity,
                    ww_batch **out);

/* Launch geometry: workgroups (4 waves each) per CU (0 = as many as fit) and the
   largest number of consecutive 8192-integer segments a wave claims at once (0 = 8);
   claims shrink toward one segment as the launch's work runs out. */
int ww_batch_set_grid(ww_batch *b, int blocks_per_cu, uint32_t claim);

/* Stream-ordered search of every range of the batch: the wide pre-sieve of its segments
   (when the context's pre-sieve bound is set; with the bucket sieve first for the ranges
   it covers), then one persistent kernel that clears the batch's sums and whose waves
   claim 8192-integer segments from a shared cursor.
   Allocates nothing and records no memset, so it may run under graph capture. */
int ww_batch_launch(ww_ctx *ctx, ww_batch *b);

/* The same launch captured once into HIP graphs on first use (one for the pre-sieve, one
   for the search) and replayed. */
int ww_batch_launch_graph(ww_ctx *ctx, ww_batch *b);

/* Blocking read: per-range sums (n entries, in table order) and up to `cap` records
   whose pad field holds the range's index; *nrec = records copied, *total = records
   produced (more than the capacity means overflow). Any pointer may be null. */
int ww_batch_read(ww_ctx *ctx, ww_batch *b, ww_range_sum *sums, ww_result *recs, uint32_t cap,
                  uint32_t *nrec, uint32_t *total);

/* Kernel milliseconds of the last completed batch launch (HIP events on the stream). */
int ww_batch_time(ww_ctx *ctx, ww_batch *b, double *ms);

/* The same launch split: the pre-sieve kernel alone and the persistent search kernel
   alone (their sum is ww_batch_time). Either pointer may be null. */
int ww_batch_kernel_times(ww_ctx *ctx, ww_batch *b, double *presieve_ms, double *search_ms);

/* The batch's bucket sieve: ranges it covers, whether the last launch used it (the
   context's sieve tiers allowed it), entries per bucket, device bytes of its buckets, and
   chunks whose buckets overflowed over all launches so far (those chunks' candidates were
   all PRP-tested: exact, only slower). */
typedef struct {
    uint32_t spans;
    uint32_t active;
    uint32_t bucket_cap;
    uint32_t overflows;
    uint64_t bytes;
} ww_bucket_info;
int ww_batch_buckets(ww_ctx *ctx, ww_batch *b, ww_bucket_info *out);

/* Waits for the batch's launches on its context's stream (not for the whole device),
   then frees it. Destroy batches before their context. */
int ww_batch_destroy(ww_batch *b);

/* One configuration of the persistent form and its measured speed. */
typedef struct {
    uint32_t deep_bound;      /* ww_set_deep_bound */
    int32_t blocks_per_cu;    /* ww_batch_set_grid; 0 = as many as fit */
    uint32_t claim;           /* largest segments per claim */
    uint32_t presieve_bound;  /* ww_set_presieve_bound */
    double ms;                /* kernel time of the sample (pre-sieve + search) */
    double prps_per_s;        /* tested primes per second on the sample */
    uint32_t waves_per_group; /* ww_set_launch_shape */
    uint32_t segment_bits;    /* ww_set_launch_shape */
} ww_tuning;

/* The autotuner's search space; a null list (or count 0) takes the built-in one:
   deep bounds {0, 1024, 2048, 4096, 8192, 16384, 65535}, workgroups per CU {0, 4, 5}
   (scaled to the shape: as many as fit, and 4 or 5 waves per SIMD), claims {1, 4, 16, 64},
   pre-sieve bounds {0, 2^18, 2^19, 2^20, 2^21}, waves per workgroup {2, 4, 8}, segment
   bits {15, ..., 20}. */
typedef struct {
    const uint32_t *deep_bounds;
    uint32_t n_deep_bounds;
    const int32_t *blocks_per_cu;
    uint32_t n_blocks_per_cu;
    const uint32_t *claims;
    uint32_t n_claims;
    const uint32_t *presieve_bounds;
    uint32_t n_presieve_bounds;
    const uint32_t *waves_per_group;
    uint32_t n_waves_per_group;
    const uint32_t *segment_bits;
    uint32_t n_segment_bits;
} ww_tune_space;

/* Autotuner replacing the reference's work-size profiling (profileGPU,
   main.cpp:959-1107): times the persistent form (pre-sieve + search kernel) on the sample
   range [lo, hi) in stages, each at the best of the ones before: deep-sieve bound x
   workgroups per CU x claim size; the work-group shape (waves per workgroup, with its
   workgroups per CU); the pre-sieve bound; the pre-sieve's segment size. space may be
   null. Keeps the fastest as this context's bounds, launch shape and the default geometry
   of batches created afterwards, and returns it in *best. `trials` (may be null) receives
   up to cap measured configurations; *ntrials their count. Results do not depend on any
   of these settings. */
int ww_autotune(ww_ctx *ctx, uint64_t lo, uint64_t hi, const ww_tune_space *space,
                ww_tuning *best, ww_tuning *trials, uint32_t cap, uint32_t *ntrials);

/* The context's tuned configuration (bounds, geometry, launch shape) as a small JSON file
   naming the device (marketing name, target, compute units), and back: ww_tuning_load
   applies a file written on the same kind of device and refuses (WW_E_ARG) one from
   another, or a malformed one. A tuning cache: the drop-in loads it at start instead of
   tuning again. No counterpart in the reference. */
int ww_tuning_save(ww_ctx *ctx, const char *path);
int ww_tuning_load(ww_ctx *ctx, const char *path, ww_tuning *applied);

/* ---------------------------------------------------------------------------
 * Serving form: continuous batching of range requests. No counterpart in the
 * reference (one range per process). An engine thread admits requests into device
 * slots as slots and HBM budget free up and runs one persistent launch per tick over
 * the next segments of every active request; per request the checksum, count and
 * records equal the plain form's for that range alone.
 * ------------------------------------------------------------------------- */

typedef struct ww_server ww_server;

typedef struct {
    uint32_t slots;             /* requests resident at once (0 = 4096) */
    uint32_t tick_segments;     /* 8192-integer segments per launch (0 = 2^17) */
    uint64_t hbm_budget;        /* device bytes for the admitted requests' resident sieve state,
                                   allocated as one arena (0 = 64 GiB; capped at 90% of free HBM) */
    uint64_t bytes_per_request; /* reserved per request on top of its sieve state (0 = none) */
    uint32_t claim;             /* segments per claim (0 = 1: shortest tick tails) */
    int32_t blocks_per_cu;      /* workgroups per CU (-1 = context default, 0 = as many as fit) */
} ww_server_config;

typedef struct {
    uint64_t id, lo, hi;
    uint64_t checksum, prp_count; /* as the plain form gives for [lo, hi) alone */
    double t_submit, t_admit, t_done; /* seconds since the server started */
    uint32_t nrecords, pad;
} ww_request_info;

typedef struct {
    double t_start;
    uint64_t submitted, completed, tested, ticks, segments, pending, active;
    uint64_t reserved_bytes;      /* HBM budget held by the active requests now */
    uint64_t peak_reserved_bytes; /* the most held at once */
    uint64_t peak_active;         /* the most requests active at once */
} ww_server_stats;

/* Start a server on the context's device (uses its deep-sieve table and tuned geometry;
   cfg may be null). Allocates device and pinned memory, including the HBM-budget arena.
   An admitted request holds, until it completes, its resident sieve state: the wide
   pre-sieve's words for all its segments (512 bytes per 8192 integers, written once at
   admission), plus bytes_per_request. Requests are admitted first come, first served,
   when a slot and a free run of the arena of their size are available. */
int ww_server_create(ww_ctx *ctx, const ww_server_config *cfg, ww_server **out);

/* Queue a request for [lo, hi) (WW_MIN_START <= lo < hi); non-blocking. WW_E_CAPACITY
   for a request whose reservation exceeds the whole HBM budget. */
int ww_server_submit(ww_server *s, uint64_t lo, uint64_t hi, uint64_t *id);

/* Completed requests, oldest first: up to cap into out, count in *n; waits up to
   wait_ms for the first one. Their records stay available to ww_server_records. */
int ww_server_poll(ww_server *s, ww_request_info *out, uint32_t cap, uint32_t *n, int wait_ms);

/* The result records of a polled request (pad = slot index); releases them. */
int ww_server_records(ww_server *s, uint64_t id, ww_result *out, uint32_t cap, uint32_t *n);

int ww_server_get_stats(ww_server *s, ww_server_stats *out);

/* Stop after the work already submitted, then free everything. */
int ww_server_destroy(ww_server *s);

/* ---------------------------------------------------------------------------
 * Allocation-free search. No counterpart in the reference: an entry point for callers
 * that cannot allocate device memory during the call (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).
 * ------------------------------------------------------------------------- */

/* Most ranges one allocation-free call takes. */
#define WW_NOALLOC_MAX_RANGES 4096u

/* The persistent search of n ranges (1 <= n <= WW_NOALLOC_MAX_RANGES) as the batched form
   does it (ww_batch_create + ww_batch_launch + ww_batch_read), but with every buffer in
   device memory the caller owns: d_ranges (n ww_range, input; a range that is not
   WW_MIN_START <= lo < hi is searched as empty), d_sums (n ww_range_sum, output: per
   range as the plain form gives it), d_records (record_cap records, pad = range index;
   order unspecified) and d_record_count (one uint32_t: records produced, which may exceed
   record_cap). It calls no hipMalloc, hipFree or allocating library routine: the span
   table, segment prefix and kernel state (about 200 KB) live in one fixed device array
   per GPU, whose address is fetched when a context is opened on that GPU. Every call on
   a GPU shares that array, so calls that overlap on one GPU (from contexts on different
   streams) race; calls on one context are ordered on its stream. It runs without the wide
   pre-sieve (same results, slower launch). Stream-ordered: ww_finish, then read the
   outputs. */
int ww_search_ranges_noalloc(ww_ctx *ctx, const ww_range *d_ranges, uint32_t n,
                             ww_range_sum *d_sums, ww_result *d_records, uint32_t record_cap,
                             uint32_t *d_record_count);

/* Measurement support, no counterpart in the reference: the issue rate of
   independent 32x32->64 multiplies on every CU of `device` (kind 0: v_mad_u64_u32,
   kind 1: v_mul_hi_u32), in operations per second. Prices the roofline of the
   path, which is bound by these multiplies. */
int ww_peak_mul32(int device, int kind, double *ops_per_s);

/* Measurement support, no counterpart in the reference: one launch of a plain copy
   kernel moving `bytes` (a multiple of 16) between two fresh device buffers; *ms is
   its duration. Calibrates the HBM traffic counters (FETCH_SIZE, WRITE_SIZE) against
   a known byte count. */
int ww_copy_probe(int device, uint64_t bytes, double *ms);

#ifdef __cplusplus
}
#endif

#endif
