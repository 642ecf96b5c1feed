/*
 * ww_batch.h -- device layout of the persistent (batched / serving) form.
 */
#ifndef WW_BATCH_H
#define WW_BATCH_H

#include "ww_kernels.h"

namespace ww {

/* Segments [seg0, seg0 + count) of the range [lo, hi) of slot `slot`; segment k covers
   the integers [(lo | 1) + 8192 k, (lo | 1) + 8192 (k + 1)) within the range. The count
   is implied by the prefix array. */
struct BatchSpan {
    uint64_t lo, hi, seg0;
    uint32_t slot, pad;
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
inline uint64_t segments_of(uint64_t lo, uint64_t hi)
{
    uint64_t base = lo | 1;
    return base >= hi ? 0 : (hi - base + kIntsPerWave - 1) / kIntsPerWave;
}

int batch_blocks_per_cu();

/* Clears the state and the first `nslots` slot sums (0: none, so sums accumulate over
   launches), then runs the persistent kernel over all spans; each wave claims `claim`
   consecutive segments at a time.
   No allocation, no memset: safe to capture into a graph. */
hipError_t launch_batch(const BatchSpan *spans, const uint64_t *prefix, uint32_t nspans,
                        BatchState *st, SlotSum *sums, uint32_t nslots, ww_result *res,
                        uint32_t res_cap, int grid, uint32_t claim, const DeepPrime *deep,
                        uint32_t ndeep, hipStream_t s);

} // namespace ww

#endif
