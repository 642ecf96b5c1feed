/*
 * ww_bucket.hip -- the bucket sieve of the persistent form: sieving primes from above the
 * wide pre-sieve's bound up to the square root of each range's end.
 *
 * With it, every odd number of a range is sieved by all primes up to sqrt(hi): by 3..113
 * in registers and 127..B in LDS (k_batch), by B..B2 per 2^20-bit chunk (k_presieve) and
 * by the primes above B2 here. Each tier drops the multiples of its primes except the
 * numbers in a prime's base-2 pseudoprime class, which it flags instead. A candidate that
 * is neither dropped nor flagged then has no prime factor up to its square root, so it is
 * prime and needs no strong PRP test; a flagged one is a multiple of a sieving prime and
 * is tested as before. The PRP list (and so every checksum) is unchanged.
 *
 * Primes above 2^20 hit a 2^20-bit chunk at most once, so computing every prime's start
 * per chunk (as k_presieve does for the smaller ones) would be wasted. Instead k_bucket_fill
 * walks each prime once over the whole range and drops its hits into buckets, one per
 * (chunk, workgroup) pair, which k_presieve then applies to the chunk's LDS bitmap. The
 * waves take blocks of 64 consecutive primes round-robin over the workgroups, so every
 * workgroup gets the same mix of small primes (thousands of hits each) and large ones (a
 * handful), and the bucket sizes stay close to their mean. Each workgroup counts its hits per chunk in LDS,
 * which also hands out the slots of its own bucket, so no global atomics are needed and the
 * bucket of (chunk, workgroup) fills sequentially.
 *
 * The sieving primes (every odd prime below 2^32, 813 MB) are generated once per context on
 * the device by k_primes32, a segmented sieve by the primes below 2^16.
 */
#include "ww_batch.h"
#include "ww_bucket.h"

namespace ww {

namespace {

constexpr int kFillThreads = 1024;
constexpr int kPrimeThreads = 1024;
constexpr u32 kPrimeBlockBits = 1u << 20; /* odd numbers per block of the prime sieve */
constexpr u32 kPrimeBlockWords = kPrimeBlockBits / 32;
constexpr u32 kPrimeWordsPerThread = kPrimeBlockWords / kPrimeThreads; /* 32 */

/* Exclusive prefix sum of one value per thread over a 1024-thread workgroup; returns the
   thread's prefix and stores the total in *total (same in every thread). */
__device__ __forceinline__ u32 block_scan_1024(u32 v, u32 *wsum, u32 *total)
{
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 incl = v;
    for (int d = 1; d < 64; d <<= 1) {
        const u32 t = __shfl_up(incl, d, 64);
        if (lane >= (u32)d) incl += t;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        u32 acc = 0;
        for (int w = 0; w < kPrimeThreads / 64; ++w) {
            const u32 t = wsum[w];
            wsum[w] = acc;
            acc += t;
        }
        wsum[kPrimeThreads / 64] = acc;
    }
    __syncthreads();
    *total = wsum[kPrimeThreads / 64];
    return wsum[wave] + incl - v;
}

/* Block b of the odd numbers below 2^32: n = 2^21 b + 2 i + 1, i < 2^20. Marks the odd
   multiples of the primes 3..65521 from their squares on, and 1. Without `offsets` it
   writes the block's prime count to counts[b]; with them it writes the block's primes in
   ascending order from out[offsets[b]] on. */
__global__ __launch_bounds__(kPrimeThreads) void k_primes32(const u32 *small, u32 nsmall,
                                                            u32 *counts, const u64 *offsets,
                                                            u32 *out)
{
    __shared__ u32 bm[kPrimeBlockWords];
    __shared__ u32 wsum[kPrimeThreads / 64 + 1];
    const u32 base = blockIdx.x * (2 * kPrimeBlockBits); /* < 2^32 */
    for (u32 w = threadIdx.x; w < kPrimeBlockWords; w += kPrimeThreads) bm[w] = 0;
    __syncthreads();
    for (u32 j = threadIdx.x; j < nsmall; j += kPrimeThreads) {
        const u32 l = small[j];
        u64 start = (u64)l * l; /* smaller multiples have a smaller prime factor */
        const u32 first = base + 1;
        if (start < first) {
            const u32 r = first % l;
            u64 m = (u64)first + (r ? l - r : 0);
            if ((m & 1) == 0) m += l;
            start = m;
        }
        for (u64 i = (start - first) >> 1; i < kPrimeBlockBits; i += l)
            atomicOr(&bm[i >> 5], 1u << (i & 31));
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&bm[0], 1u); /* 1 is not prime */
    __syncthreads();
    const u32 w0 = threadIdx.x * kPrimeWordsPerThread;
    u32 c = 0;
    for (u32 k = 0; k < kPrimeWordsPerThread; ++k) c += (u32)__popc(~bm[w0 + k]);
    u32 total;
    const u32 pre = block_scan_1024(c, wsum, &total);
    if (!offsets) {
        if (threadIdx.x == 0) counts[blockIdx.x] = total;
        return;
    }
    u64 at = offsets[blockIdx.x] + pre;
    for (u32 k = 0; k < kPrimeWordsPerThread; ++k) {
        u32 v = ~bm[w0 + k];
        while (v) {
            const u32 i = (w0 + k) * 32 + (u32)__builtin_ctz(v);
            out[at++] = base + 2 * i + 1;
            v &= v - 1;
        }
    }
}

/* The bucket fill. For every bucket span f, every workgroup walks its share of the
   span's sieving primes (blocks of 64 consecutive table entries, round-robin over the
   workgroups, up to the first prime above the span's top) and writes each hit
   to its (chunk, workgroup) bucket: entry = odd-number index in the chunk, top bit set
   when the hit is in the prime's pseudoprime class. The counts are written for every
   (chunk, workgroup) pair of the span, so nothing needs clearing between launches; a count
   above `cap` means entries were lost and k_presieve treats that chunk as not sieved. */
__global__ __launch_bounds__(kFillThreads) void k_bucket_fill(const BatchSpan *spans,
                                                              const u64 *prefix,
                                                              const BucketSpan *bspans,
                                                              u32 nb, const u32 *primes,
                                                              u32 *entries, u32 *counts,
                                                              u32 cap)
{
    __shared__ u32 cnt[kBucketMaxChunks];
    __shared__ uint16_t res[kWheelRes];
    __shared__ uint8_t gap[kWheelRes];
    __shared__ uint8_t nxt[kWheelMod];
    {
        constexpr WheelTables wt = make_wheel();
        const u32 t = threadIdx.x;
        if (t < kWheelRes) {
            res[t] = wt.res[t];
            gap[t] = wt.gap[t];
        }
        if (t < kWheelMod) nxt[t] = wt.next[t];
    }
    const u32 groups = gridDim.x, wg = blockIdx.x;
    for (u32 f = 0; f < nb; ++f) {
        const BucketSpan bs = bspans[f];
        const BatchSpan sp = spans[bs.span];
        const u64 G0 = (sp.lo | 1) + sp.seg0 * kIntsPerWave;
        const u64 nbits = (prefix[bs.span + 1] - prefix[bs.span]) * 4096;
        for (u32 j = threadIdx.x; j < bs.nchunks; j += kFillThreads) cnt[j] = 0;
        __syncthreads();
        /* wave v of workgroup w takes the 64-prime blocks w + groups (v + 16 r), r = 0, 1,
           ...: neighbouring workgroups get neighbouring blocks, so their hit counts per
           chunk differ by about one block's, and each wave reads 256 contiguous bytes */
        const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (u64 blk = bs.pidx0 + 64 * ((u64)wg + (u64)groups * wave); blk < bs.pidx1;
             blk += 64 * (u64)groups * (kFillThreads / 64)) {
            if (primes[blk] > bs.top) break; /* ascending: the rest is above the top too */
            const u64 i = blk + lane;
            if (i >= bs.pidx1) continue;
            const u32 l = primes[i];
            if (l > bs.top) continue;
            const double linv = 1.0 / (double)l;
            BucketStart s = bucket_start(G0, l, linv, res, nxt);
            u64 b = s.b;
            u32 idx = s.idx, x = s.x;
            while (b < nbits) {
                const u64 B = bs.bit0 + b;
                const u32 j = (u32)(B >> kBucketChunkBits) - bs.chunk_first;
                const u32 slot = atomicAdd(&cnt[j], 1u);
                if (slot < cap)
                    entries[bs.ent_base + ((u64)j * groups + wg) * cap + slot] =
                        ((u32)B & ((1u << kBucketChunkBits) - 1)) | (x == 1u ? kBucketRescue : 0u);
                bucket_step(b, idx, x, l, linv, gap);
            }
        }
        __syncthreads();
        for (u32 j = threadIdx.x; j < bs.nchunks; j += kFillThreads)
            counts[bs.cnt_base + (u64)j * groups + wg] = cnt[j];
        __syncthreads();
    }
}

} // namespace

hipError_t launch_primes32(const u32 *small, u32 nsmall, u32 nblocks, u32 *counts,
                           const u64 *offsets, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_primes32, dim3(nblocks), dim3(kPrimeThreads), 0, s, small, nsmall,
                       counts, offsets, out);
    return hipGetLastError();
}

hipError_t launch_bucket_fill(const BatchSpan *spans, const u64 *prefix, const BucketSpan *bspans,
                              u32 nb, const u32 *primes, u32 *entries, u32 *counts, u32 cap,
                              u32 groups, hipStream_t s)
{
    hipLaunchKernelGGL(k_bucket_fill, dim3(groups), dim3(kFillThreads), 0, s, spans, prefix,
                       bspans, nb, primes, entries, counts, cap);
    return hipGetLastError();
}

} // namespace ww
