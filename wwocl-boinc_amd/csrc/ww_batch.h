/*
 * ww_batch.h -- device layout of the persistent (batched / serving) form.
 */
#ifndef WW_BATCH_H
#define WW_BATCH_H

#include "ww_kernels.h"

namespace ww {

/* Segments [seg0, seg0 + count) of the range [lo, hi) of slot `slot`; segment k covers
   the integers [(lo | 1) + 8192 k, (lo | 1) + 8192 (k + 1)) within the range. The count
   is implied by the prefix array. The span's first segment reads its pre-sieve words at
   block pre_seg of the pre-sieve buffer (64 words per block): the span's position in the
   launch (its prefix) when the buffer follows the launch's layout, or where the range's
   resident words start (the serving form). */
struct BatchSpan {
    uint64_t lo, hi, seg0;
    uint32_t slot, pad;
    uint64_t pre_seg;
};

/* Per range slot: checksum and number of tested primes (list entries). */
struct SlotSum {
    uint64_t checksum;
    uint64_t count;
};

struct BatchState {
    uint64_t cursor;      /* next segment to claim */
    uint32_t resultcount; /* records produced (may exceed capacity) */
    uint32_t reserved;
};

/* Number of segments of [lo, hi) (lo | 1 < hi). */
__host__ __device__ inline uint64_t segments_of(uint64_t lo, uint64_t hi)
{
    uint64_t base = lo | 1;
    return base >= hi ? 0 : (hi - base + kIntsPerWave - 1) / kIntsPerWave;
}

/* One prime of the wide pre-sieve; same layout as ww_wide_prime of the C ABI. */
struct WidePrime {
    uint32_t ell, pad;
    uint64_t m, bar_ell, bar_m;
};

/* One range of a batch that the bucket sieve covers (sieving primes above the wide
   pre-sieve, up to sqrt(hi - 1)); its BatchSpan's pad holds 1 + its index in this table.
   The fill writes each hit to a band (kBucketBandChunks consecutive 2^20-bit chunks) of
   its workgroup: workgroup w, band b holds up to band_cap entries from
   band_entries[band_ent_base + (w nbands + b) band_cap] on (a workgroup's buckets are
   contiguous), its count at band_counts[band_cnt_base + w nbands + b]. The split then
   gathers each chunk's entries in kBucketSplitParts parts (one per split workgroup of the
   band): chunk j, part p holds up to cap entries from entries[ent_base + (j parts + p) cap]
   on, its count at counts[cnt_base + j parts + p]. */
struct BucketSpan {
    uint64_t ent_base, cnt_base;           /* per (chunk, part) */
    uint64_t band_ent_base, band_cnt_base; /* per (fill workgroup, band) */
    uint64_t pidx0, pidx1; /* sieving primes: table entries [pidx0, pidx1) up to top */
    uint64_t bit0;         /* global odd-number index of the span's first segment */
    uint32_t span;         /* index of the span in the span table */
    uint32_t chunk_first;  /* bit0 >> 20 */
    uint32_t nchunks;      /* chunks holding the span's segments */
    uint32_t top;          /* isqrt(hi - 1): the largest prime to sieve by */
    uint32_t nbands;       /* ceil(nchunks / kBucketBandChunks) */
    uint32_t pad;
};
/* Chunks one bucket span may cover (a 2^34-wide range and one boundary chunk). */
constexpr uint32_t kBucketMaxChunks = 8200;
/* Chunks per band: a fill workgroup keeps one partly written line per band open, and 2^33
   bits make 256 bands, so the open lines of the workgroups of one XCD fit in its L2. */
constexpr uint32_t kBucketBandChunks = 16;
constexpr uint32_t kBucketMaxBands = (kBucketMaxChunks + kBucketBandChunks - 1) / kBucketBandChunks;
/* Workgroups of the bucket fill (one per CU: half the open lines of two per CU, 2.93 ->
   2.68 ms per 2^33 range), and so band buckets per band. */
constexpr uint32_t kBucketGroups = 256;
/* Split workgroups per band, each taking an equal share of the fill workgroups' buckets;
   and so buckets per chunk. */
constexpr uint32_t kBucketSplitParts = 4;

/* Words of the pre-sieve per segment (one per lane of a wave). */
constexpr uint64_t kPresieveWordsPerSeg = 64;
/* The pre-sieve skips primes up to this (their pseudoprime classes are too dense) ... */
constexpr uint32_t kPresieveMinPrime = 1024;
/* ... and pieces of fewer segments than this, where computing every prime's start
   positions would cost more than the strong PRP tests it saves (those segments drop
   nothing extra). */
constexpr uint64_t kPresieveMinSegs = 16;
/* The pre-sieve's LDS segment, in 8192-integer segments: a power of two from 8 (2^15 bits,
   one per odd number) to 256 (2^20 bits, the default). */
constexpr uint32_t kPresieveMinChunkSegs = 8;
constexpr uint32_t kPresieveMaxChunkSegs = 256;
bool presieve_chunk_ok(uint32_t chunk_segs);

/* Work-group shapes of the persistent kernel: 2, 4 or 8 waves per workgroup. */
constexpr int kDefaultBatchWaves = 4;
bool batch_waves_ok(int waves);
/* Resident workgroups per CU of the persistent kernel at that shape (0 on error). */
int batch_blocks_per_cu(int waves = kDefaultBatchWaves);

/* Buckets of a launch for k_presieve: the bucket spans, their entries and counts, the
   entries per bucket, and a counter of chunks whose buckets overflowed. */
struct BucketArgs {
    const BucketSpan *spans;
    const uint32_t *entries;
    const uint32_t *counts;
    uint32_t cap;
    uint32_t *overflows;
};

/* The wide pre-sieve over all segments of the spans: out[64 s + w] = odd numbers to drop
   in word w of segment s (multiples of tab's primes, minus their pseudoprime classes).
   With `outr` it also writes outr[64 s + w] = the odd numbers that still need the strong
   PRP test: those in a pseudoprime class of a sieving prime, and all of them in pieces
   that are not sieved up to sqrt(hi) (no bucket span, short pieces, overflowed lists);
   the pieces of bucket spans then also get the buckets of `bk` applied (chunk_segs must
   be 256). No allocation, no memset: safe to capture into a graph. */
hipError_t launch_presieve(const BatchSpan *spans, const uint64_t *prefix, uint32_t nspans,
                           const WidePrime *tab, uint32_t ntab, uint64_t *out, int grid,
                           hipStream_t s, uint32_t chunk_segs = kPresieveMaxChunkSegs,
                           uint64_t *outr = nullptr, const BucketArgs *bk = nullptr);

/* Every odd prime below 2^32 into out (203,280,220 entries), ascending: one launch with
   offsets = null writes the prime count of each 2^21-integer block to counts, a second
   with their exclusive prefix sums in offsets writes the primes. small = the odd primes
   below 2^16. */
hipError_t launch_primes32(const uint32_t *small, uint32_t nsmall, uint32_t nblocks,
                           uint32_t *counts, const uint64_t *offsets, uint32_t *out,
                           hipStream_t s);

/* The bucket sieve of the bucket spans (see BucketSpan): the fill (kBucketGroups
   workgroups) into bands, then the split into per-chunk buckets (kBucketSplitParts
   workgroups per band, `nbands` bands in all). */
struct BucketStore {
    uint32_t *band_entries, *band_counts;
    uint32_t band_cap;
    uint32_t *entries, *counts;
    uint32_t cap;
};
hipError_t launch_bucket_sieve(const BatchSpan *spans, const uint64_t *prefix,
                               const BucketSpan *bspans, uint32_t nb, uint32_t nbands,
                               const uint32_t *primes, const uint32_t *orders,
                               const BucketStore &st, hipStream_t s);

/* orders[i] = ord_l(2) for every table prime l = primes[i] above 2^20, 0 for the others
   (i < n, the whole table of odd primes below 2^32); small = the odd primes below 2^16 with
   their divisibility data (ww_bucket.h's SmallDiv, nsmall of them); bounds = device
   scratch for 2^16 + 1 table indices. */
struct SmallDiv;
hipError_t launch_order2(const uint32_t *primes, uint64_t n, const SmallDiv *small,
                         uint32_t nsmall, uint32_t *orders, uint64_t *bounds, hipStream_t s);

/* Writes the one-span table of the range [lo, hi) (segments from lo | 1) into span and
   prefix (device scratch), so that a single range needs no host copy. */
hipError_t launch_set_span(uint64_t lo, uint64_t hi, BatchSpan *span, uint64_t *prefix,
                           hipStream_t s);

/* launch_set_span for a fused chunk with the bucket sieve: the span's pad = 1, and bs
   (bucket span 0 of that one-span table) written to *bspan. */
hipError_t launch_set_bucket_span(uint64_t lo, uint64_t hi, BatchSpan *span, uint64_t *prefix,
                                  BucketSpan *bspan, const BucketSpan &bs, hipStream_t s);

/* launch_set_span, then launch_presieve over that one span (a plain-form chunk). */
hipError_t launch_presieve_range(uint64_t lo, uint64_t hi, BatchSpan *span, uint64_t *prefix,
                                 const WidePrime *tab, uint32_t ntab, uint64_t *out, int grid,
                                 hipStream_t s, uint32_t chunk_segs = kPresieveMaxChunkSegs);

/* Clears the state and the first `nslots` slot sums (0: none, so sums accumulate over
   launches), then runs the persistent kernel over all spans; each wave claims up to
   `claim` consecutive segments at a time. `pre` (may be null) is the wide pre-sieve of
   the same spans. `waves` is the work-group shape (2, 4 or 8 waves). `prer` (may be
   null: every candidate is tested) is the pre-sieve's words of candidates that need the
   strong PRP test; the others are known primes. It requires the deep sieve up to at least
   1024 (the pre-sieve's lowest prime), whose pseudoprime classes it flags as well.
   No allocation, no memset: safe to capture into a graph. */
hipError_t launch_batch(const BatchSpan *spans, const uint64_t *prefix, uint32_t nspans,
                        BatchState *st, SlotSum *sums, uint32_t nslots, ww_result *res,
                        uint32_t res_cap, int grid, uint32_t claim, const DeepPrime *deep,
                        uint32_t ndeep, const uint64_t *pre, hipStream_t s,
                        int waves = kDefaultBatchWaves, const uint64_t *prer = nullptr);

/* Adds one single-range persistent launch to the plain form's counters and records: its
   checksum and tested count to the totals, and its records (up to bcap) after the ones
   already held (pad cleared). The PRP list and its count are left alone. */
hipError_t launch_fold(const BatchState *bst, const SlotSum *sum, const ww_result *bres,
                       uint32_t bcap, DevState *ds, ww_result *res, uint32_t cap, hipStream_t s);

/* The allocation-free search (ww_noalloc.hip): noalloc_init fetches this device's
   address of the fixed scratch array (once per device, with the device current);
   launch_noalloc builds the spans and prefix there from the caller's ranges, clears the
   caller's sums, runs the persistent kernel and writes the record count. */
hipError_t noalloc_init(int device);
hipError_t launch_noalloc(int device, const ww_range *ranges, uint32_t n, SlotSum *sums,
                          ww_result *res, uint32_t res_cap, uint32_t *res_count, int grid,
                          uint32_t claim, const DeepPrime *deep, uint32_t ndeep, hipStream_t s,
                          int waves = kDefaultBatchWaves);

} // namespace ww

#endif
