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
 * (band of 16 chunks, workgroup) pair; k_bucket_split sorts each band's entries into
 * per-chunk buckets, which k_presieve then applies to the chunk's LDS bitmap. The fill's
 * waves take blocks of 64 consecutive primes round-robin over the workgroups, so every
 * workgroup gets the same mix of small primes (thousands of hits each) and large ones (a
 * handful), and the bucket sizes stay close to their mean. Each workgroup counts its hits
 * per band in LDS, which also hands out the slots of its own bucket, so no global atomics
 * are needed and the bucket of (band, workgroup) fills sequentially.
 *
 * The sieving primes (every odd prime below 2^32, 813 MB) are generated once per context on
 * the device by k_primes32, a segmented sieve by the primes below 2^16, and their base-2
 * orders (another 813 MB) by k_order2.
 */
#include "ww_batch.h"
#include "ww_bucket.h"

namespace ww {

namespace {

constexpr int kFillThreads = 1024;

/* A bucket span has at most kBucketMaxChunks chunks of 2^20 odd numbers and its sieving
   primes are above 2^20, so a walk meets at most 2 kBucketMaxChunks + 1 cofactors: its
   wheel index stays inside the twice-stored gaps. */
static_assert(2 * (u64)kBucketMaxChunks + 1 <= kWalkMaxK, "a bucket walk's window");
constexpr int kPrimeThreads = 1024;
constexpr u32 kPrimeBlockBits = 1u << 20; /* odd numbers per block of the prime sieve */
constexpr u32 kPrimeBlockWords = kPrimeBlockBits / 32;
constexpr u32 kPrimeWordsPerThread = kPrimeBlockWords / kPrimeThreads; /* 32 */

/* A slot in the LDS counter cnt[key] for every active lane, as atomicAdd(&cnt[key], 1)
   would give it. The lanes sharing the lowest active lane's key (all of them when a wave's
   hits cluster, as a small prime's do) take theirs with one atomic; the others add their
   own, so no 64-way conflict on one counter serialises the wave. */
__device__ __forceinline__ u32 wave_slot(u32 *cnt, u32 key)
{
    const u32 k0 = __builtin_amdgcn_readfirstlane(key);
    const bool same = key == k0;
    const u64 m = __ballot(same);
    const u32 below = __builtin_amdgcn_mbcnt_hi((u32)(m >> 32), __builtin_amdgcn_mbcnt_lo((u32)m, 0u));
    u32 base = 0;
    if (same && below == 0) base = atomicAdd(&cnt[k0], (u32)__popcll(m));
    base = __builtin_amdgcn_readfirstlane(base); /* the lowest active lane took it */
    return same ? base + below : atomicAdd(&cnt[key], 1u);
}

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

/* Index in a block (odd numbers first + 2 i) of the first odd multiple of the prime l
   that the block marks: l^2 if that is in or past the block (smaller multiples have a
   smaller prime factor), else the first odd multiple >= first. */
__device__ __forceinline__ u64 prime_block_start(u32 l, u32 first)
{
    u64 start = (u64)l * l;
    if (start < first) {
        const u32 r = first % l;
        u64 m = (u64)first + (r ? l - r : 0);
        if ((m & 1) == 0) m += l;
        start = m;
    }
    return (start - first) >> 1;
}

/* Primes below this many mark a block together (every thread a share of each one's
   multiples); the larger ones, one per thread, have at most 2^20 / 2048 = 512 marks each,
   so no thread waits on a long loop of its own. */
constexpr u32 kPrimeShared = 2048;

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
    const u32 first = base + 1;
    for (u32 w = threadIdx.x; w < kPrimeBlockWords; w += kPrimeThreads) bm[w] = 0;
    __syncthreads();
    u32 j = 0;
    for (; j < nsmall && small[j] < kPrimeShared; ++j) { /* uniform: the whole block */
        const u32 l = small[j];
        for (u64 i = prime_block_start(l, first) + (u64)threadIdx.x * l; i < kPrimeBlockBits;
             i += (u64)kPrimeThreads * l)
            atomicOr(&bm[i >> 5], 1u << (i & 31));
    }
    for (j += threadIdx.x; j < nsmall; j += kPrimeThreads) {
        const u32 l = small[j];
        for (u64 i = prime_block_start(l, first); i < kPrimeBlockBits; i += l)
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
   workgroups, up to the first prime above the span's top) and writes each hit to its
   workgroup's bucket of the hit's band: entry = odd-number index in the chunk, the chunk's
   index in the band above it, and the top bit when the hit is in the prime's pseudoprime
   class. The slots come from the workgroup's LDS cursor per band, so a band bucket fills
   front to back and only one partly written line per band is open at a time. The counts
   are written for every (band, workgroup) pair of the span, so nothing needs clearing
   between launches; a count above band_cap means entries were lost (the split passes that
   on to the band's chunks). */
__global__ __launch_bounds__(kFillThreads) void k_bucket_fill(const BatchSpan *spans,
                                                              const u64 *prefix,
                                                              const BucketSpan *bspans,
                                                              u32 nb, const u32 *primes,
                                                              const u32 *orders, BucketStore st)
{
    __shared__ u32 cur[kBucketMaxBands];
    __shared__ uint16_t res[kWheelRes];
    __shared__ uint8_t gap[2 * kWheelRes];
    __shared__ u32 wbits[kWheelWords];
    __shared__ uint16_t wrank[kWheelWords];
    __shared__ u32 visit_a[kVisitWordsA];
    __shared__ u32 visit_b[kVisitWordsB];
    __shared__ u32 visit_c[kVisitWordsC];
    {
        static constexpr WheelTables wt = make_wheel();
        for (u32 t = threadIdx.x; t < kWheelRes; t += kFillThreads) res[t] = wt.res[t];
        for (u32 t = threadIdx.x; t < 2 * kWheelRes; t += kFillThreads) gap[t] = wt.gap[t];
        for (u32 t = threadIdx.x; t < kWheelWords; t += kFillThreads) {
            wbits[t] = wt.bits[t];
            wrank[t] = wt.rank[t];
        }
        for (u32 t = threadIdx.x; t < kVisitWordsA; t += kFillThreads) visit_a[t] = wt.visit_a[t];
        for (u32 t = threadIdx.x; t < kVisitWordsB; t += kFillThreads) visit_b[t] = wt.visit_b[t];
        for (u32 t = threadIdx.x; t < kVisitWordsC; t += kFillThreads) visit_c[t] = wt.visit_c[t];
    }
    const u32 groups = gridDim.x, wg = blockIdx.x;
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (u32 f = 0; f < nb; ++f) {
        const BucketSpan bs = bspans[f];
        const BatchSpan sp = spans[bs.span];
        const u64 G0 = (sp.lo | 1) + sp.seg0 * kIntsPerWave;
        const u64 nbits = (prefix[bs.span + 1] - prefix[bs.span]) * 4096;
        for (u32 j = threadIdx.x; j < bs.nbands; j += kFillThreads) cur[j] = 0;
        const u64 wbase = bs.band_ent_base + (u64)wg * bs.nbands * st.band_cap;
        const u32 bstride = st.band_cap;
        __syncthreads();
        /* wave v of workgroup w takes the 64-prime blocks w + groups (v + 16 r), r = 0, 1,
           ...: neighbouring workgroups get neighbouring blocks, so their hit counts per
           band differ by about one block's, and each wave reads 256 contiguous bytes */
        const u64 stride = 64 * (u64)groups * (kFillThreads / 64);
        u64 blk = bs.pidx0 + 64 * ((u64)wg + (u64)groups * wave);
        /* the block's prime and order entries are loaded one block ahead, so their latency
           hides behind the walk of the block before (lanes past the span's end read its
           last entry and sit out) */
        u32 lnext = 0, onext = 0;
        if (blk < bs.pidx1) {
            const u64 ic = blk + lane < bs.pidx1 ? blk + lane : bs.pidx1 - 1;
            lnext = primes[ic];
            onext = orders[ic];
        }
        for (; blk < bs.pidx1; blk += stride) {
            const u64 i = blk + lane;
            const u32 l = lnext, ord = onext;
            /* ascending: once lane 0's prime (blk < pidx1) is above the top, so is the rest */
            if ((u32)__builtin_amdgcn_readlane((int)l, 0) > bs.top) break;
            if (blk + stride < bs.pidx1) {
                const u64 in = blk + stride + lane;
                const u64 ic = in < bs.pidx1 ? in : bs.pidx1 - 1;
                lnext = primes[ic];
                onext = orders[ic];
            }
            if (i >= bs.pidx1 || l > bs.top) continue;
            BucketWalk w = bucket_start(G0, l, inv_below(l), ord, res, wbits, wrank);
            /* walk in global distances (twice the odd-number index); band buckets of this
               workgroup at wbase + band * bstride. The next step's gap and visit words are
               read one step ahead (bucket_step and bucket_visit, pipelined), so their LDS
               latency hides behind the current step. */
            w.d += 2 * bs.bit0;
            const u64 Dend = 2 * (bs.bit0 + nbits);
            u32 g = gap[w.idx];
            u32 va = visit_a[w.ma >> 5], vb = visit_b[w.mb >> 5], vc = visit_c[w.mc >> 5];
            while (w.d < Dend) {
                const u32 gn = gap[w.idx + 1]; /* below 2 x 5760: the span's window */
                u32 man = w.ma + g, mbn = w.mb + g, mcn = w.mc + g;
                man = man >= kVisitModA ? man - kVisitModA : man;
                mbn = mbn >= kVisitModB ? mbn - kVisitModB : mbn;
                mcn = mcn >= kVisitModC ? mcn - kVisitModC : mcn;
                const u32 van = visit_a[man >> 5], vbn = visit_b[mbn >> 5], vcn = visit_c[mcn >> 5];
                if ((va >> (w.ma & 31u)) & (vb >> (w.mb & 31u)) & (vc >> (w.mc & 31u)) & 1u) {
                    const u32 j = (u32)(w.d >> (kBucketChunkBits + 1)) - bs.chunk_first;
                    const u32 band = j / kBucketBandChunks;
                    const u32 slot = wave_slot(cur, band);
                    if (slot < st.band_cap)
                        st.band_entries[wbase + (u64)band * bstride + slot] =
                            ((u32)(w.d >> 1) & ((1u << kBucketChunkBits) - 1)) |
                            ((j % kBucketBandChunks) << kBucketChunkBits) |
                            (w.c == 0u ? kBucketRescue : 0u);
                }
                bucket_advance(w, l, ord, g);
                g = gn;
                va = van;
                vb = vbn;
                vc = vcn;
            }
        }
        __syncthreads();
        for (u32 j = threadIdx.x; j < bs.nbands; j += kFillThreads)
            st.band_counts[bs.band_cnt_base + (u64)wg * bs.nbands + j] = cur[j];
        __syncthreads();
    }
}

/* The split: kBucketSplitParts workgroups per band of a bucket span, part p reading the
   band's buckets of fill workgroups p, p + parts, ... and appending their entries to its
   own bucket of each chunk of the band. A wave works on its own: it loads 16 entries per
   lane (16 loads in flight), and sorts them 64 at a time into 16 LDS bins of its own, one
   per chunk (slot from a per-bin LDS counter). A bin that reaches 64 entries is written out
   as one coalesced 256-byte store at a slot block taken from the workgroup's counter of
   that chunk, and its rest moves to the front; the partial bins are written at the end. A
   chunk part whose entries do not all fit, or whose band lost entries in the fill, gets a
   count above cap. */
constexpr u32 kSplitWaves = kFillThreads / 64;
constexpr u32 kSplitBin = 128; /* fewer than 64 left after a flush, plus 64 per step */

__device__ __forceinline__ void split_flush(u32 *bin, u32 *bcnt, u32 jj, u32 n, u32 *cnt,
                                            u32 *out, u32 cap, u32 lane)
{
    /* write the first n (<= 64) entries of bin jj to the chunk's bucket, move the rest of
       the bin (bcnt[jj] - n, below 64) to the front */
    const u32 have = bcnt[jj];
    u32 base = 0;
    if (lane == 0) base = atomicAdd(&cnt[jj], n);
    base = __builtin_amdgcn_readfirstlane(base);
    const u32 v = lane < n ? bin[jj * kSplitBin + lane] : 0u;
    const u32 rest = have - n;
    const u32 mv = lane < rest ? bin[jj * kSplitBin + n + lane] : 0u;
    if (lane < n && base + lane < cap) out[(u64)jj * kBucketSplitParts * cap + base + lane] = v;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < rest) bin[jj * kSplitBin + lane] = mv;
    if (lane == 0) bcnt[jj] = rest;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(kFillThreads) void k_bucket_split(const BucketSpan *bspans, u32 nb,
                                                               BucketStore st, u32 groups)
{
    __shared__ u32 cnt[kBucketBandChunks];
    __shared__ u32 lost;
    __shared__ u32 bins[kSplitWaves][kBucketBandChunks * kSplitBin];
    __shared__ u32 bcnts[kSplitWaves][kBucketBandChunks];
    constexpr int kUnroll = 16;
    /* this workgroup's band and part: the bands of the spans in table order */
    const u32 part = blockIdx.x % kBucketSplitParts;
    u32 f = 0, band = blockIdx.x / kBucketSplitParts;
    while (f < nb && band >= bspans[f].nbands) band -= bspans[f++].nbands;
    if (f >= nb) return;
    const BucketSpan bs = bspans[f];
    const u32 lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    u32 *bin = bins[wave];
    u32 *bcnt = bcnts[wave];
    if (threadIdx.x < kBucketBandChunks) cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) lost = 0;
    if (lane < kBucketBandChunks) bcnt[lane] = 0;
    __syncthreads();
    const u64 first_chunk = (u64)band * kBucketBandChunks;
    /* chunk jj's bucket of this part: out + jj * parts * cap */
    u32 *out = st.entries + bs.ent_base + (first_chunk * kBucketSplitParts + part) * st.cap;
    for (u32 w = part + kBucketSplitParts * wave; w < groups; w += kBucketSplitParts * kSplitWaves) {
        u32 n = __builtin_amdgcn_readfirstlane(st.band_counts[bs.band_cnt_base + (u64)w * bs.nbands + band]);
        if (n > st.band_cap) {
            if (lane == 0) lost = 1;
            n = st.band_cap;
        }
        const u32 *run = st.band_entries + bs.band_ent_base + ((u64)w * bs.nbands + band) * st.band_cap;
        for (u32 i0 = 0; i0 < n; i0 += 64 * kUnroll) {
            u32 e[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const u32 i = i0 + 64 * u + lane;
                e[u] = i < n ? run[i] : 0xffffffffu;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (i0 + 64 * u >= n) break; /* wave-uniform */
                if (e[u] != 0xffffffffu) {
                    const u32 jj = (e[u] >> kBucketChunkBits) & (kBucketBandChunks - 1);
                    const u32 slot = wave_slot(bcnt, jj);
                    bin[jj * kSplitBin + slot] = e[u] & (((1u << kBucketChunkBits) - 1) | kBucketRescue);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                /* bins that reached 64 entries */
                u64 full = __ballot(lane < kBucketBandChunks && bcnt[lane] >= 64);
                while (full) {
                    const u32 jj = (u32)__builtin_ctzll(full);
                    split_flush(bin, bcnt, jj, 64, cnt, out, st.cap, lane);
                    full &= full - 1;
                }
            }
        }
    }
    /* the partial bins */
    for (u32 jj = 0; jj < kBucketBandChunks; ++jj) {
        const u32 have = __builtin_amdgcn_readfirstlane(bcnt[jj]);
        if (have) split_flush(bin, bcnt, jj, have, cnt, out, st.cap, lane);
    }
    __syncthreads();
    if (threadIdx.x < kBucketBandChunks && first_chunk + threadIdx.x < bs.nchunks)
        st.counts[bs.cnt_base + (first_chunk + threadIdx.x) * kBucketSplitParts + part] =
            lost ? st.cap + 1 : cnt[threadIdx.x];
}

/* The order table: orders[i] = ord_l(2) for every table prime l = primes[i] above 2^20 (0
   for the smaller ones), one workgroup per block of 2^16 integers. The odd prime factors
   of l - 1 come in two parts:
   - those above 2^11 by a sieve: for each odd prime q in (2^11, 2^16) the workgroup walks
     the odd n of its block with q | n - 1 and records q's table index in an LDS slot per
     odd n. At most two such q divide a number below 2^32 (three would exceed 2^33), so two
     16-bit halves of one word hold them;
   - those below 2^11 by trial division, one loop per wave over the 308 odd primes below
     2^11 (scalar loads, eight at a time), ending once q^2 exceeds every lane's unfactored
     part m. A lane that goes on past its own bound cannot find a new factor.
   What is left of m after both is 1 or a prime above 2^16 (every composite below 2^32 has
   a factor below 2^16), or a prime below 2^11 when the loop ended early. The factors go to
   register slots, and order2_from_factors (shared with the host) finishes. */
constexpr int kOrderThreads = 1024;
constexpr u32 kOrderBlockBits = 16;
constexpr u32 kOrderBlockOdd = 1u << (kOrderBlockBits - 1);
constexpr u32 kOrderTrial = 308; /* odd primes below 2^11: the first table entries */

/* bounds[t] = the first table index whose prime is at least t 2^16, t < nbounds. */
__global__ void k_order_bounds(const u32 *primes, u64 n, u32 nbounds, u64 *bounds)
{
    const u32 t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nbounds) return;
    const u64 v = (u64)t << kOrderBlockBits;
    u64 a = 0, b = n;
    while (a < b) {
        const u64 mid = (a + b) >> 1;
        if ((u64)primes[mid] < v) a = mid + 1;
        else b = mid;
    }
    bounds[t] = a;
}

__device__ __forceinline__ void append_factor(u32 *f, u32 &nf, u32 q)
{
#pragma unroll
    for (u32 k = 0; k < kOrderFactors; ++k)
        if (k == nf) f[k] = q;
    ++nf;
}

__global__ __launch_bounds__(kOrderThreads) void k_order2(const u32 *primes, const u64 *bounds,
                                                          const SmallDiv *__restrict__ tab,
                                                          u32 ntab, u32 *orders)
{
    __shared__ u32 big[kOrderBlockOdd]; /* per odd n of the block: 1 + two table indices */
    const u32 blk = blockIdx.x;
    const u64 i0 = bounds[blk], i1 = bounds[blk + 1];
    const u32 base = blk << kOrderBlockBits;
    if (base < (1u << kBucketChunkBits)) { /* no orders below the bucket tier */
        for (u64 i = i0 + threadIdx.x; i < i1; i += kOrderThreads) orders[i] = 0;
        return;
    }
    for (u32 w = threadIdx.x; w < kOrderBlockOdd; w += kOrderThreads) big[w] = 0;
    __syncthreads();
    /* odd n = base + 1 + 2 x: q | n - 1 iff x = -base/2 (mod q) */
    for (u32 j = kOrderTrial + threadIdx.x; j < ntab; j += kOrderThreads) {
        const u32 q = tab[j].q;
        const u32 r = (base >> 1) % q;
        for (u32 x = r ? q - r : 0u; x < kOrderBlockOdd; x += q) {
            const u32 v = j + 1;
            if (atomicCAS(&big[x], 0u, v) != 0u) atomicOr(&big[x], v << 16);
        }
    }
    __syncthreads();
    /* the block's primes, one lane each (the lanes of a wave stay together) */
    const u64 count = i1 - i0;
    for (u64 t0 = 0; t0 < count; t0 += kOrderThreads) {
        const u64 i = i0 + t0 + threadIdx.x;
        const bool act = t0 + threadIdx.x < count;
        const u32 l = act ? primes[i] : 3u;
        u32 m = act ? (l - 1) >> __builtin_ctz(l - 1) : 1u;
        u32 f[kOrderFactors];
        u32 nf = 0;
#pragma unroll
        for (u32 k = 0; k < kOrderFactors; ++k) f[k] = 0;
        constexpr u32 kGroup = 8;
        for (u32 j0 = 0; j0 < kOrderTrial; j0 += kGroup) {
            if (__ballot(tab[j0].q * tab[j0].q <= m) == 0) break;
            SmallDiv g[kGroup];
#pragma unroll
            for (u32 u = 0; u < kGroup; ++u) g[u] = tab[j0 + u < kOrderTrial ? j0 + u : kOrderTrial - 1];
#pragma unroll
            for (u32 u = 0; u < kGroup; ++u) {
                const SmallDiv d = g[u];
                if (j0 + u < kOrderTrial && small_divides(m, d)) {
                    append_factor(f, nf, d.q);
                    do m *= d.qinv; while (small_divides(m, d));
                }
            }
        }
        if (act) {
            const u32 bv = big[(l - base - 1) >> 1];
#pragma unroll
            for (u32 h = 0; h < 2; ++h) {
                const u32 v = (bv >> (16 * h)) & 0xffffu;
                if (v) {
                    const SmallDiv d = tab[v - 1];
                    append_factor(f, nf, d.q);
                    do m *= d.qinv; while (small_divides(m, d));
                }
            }
        }
        if (m > 1u) append_factor(f, nf, m);
        if (act) orders[i] = order2_from_factors(l, inv_below(l), f, nf);
    }
}

} // namespace

hipError_t launch_order2(const u32 *primes, u64 n, const SmallDiv *small, u32 nsmall,
                         u32 *orders, u64 *bounds, hipStream_t s)
{
    /* blocks of 2^16 integers up to 2^32, and their table bounds (2^16 + 1 of them) */
    constexpr u32 kBlocks = 1u << (32 - kOrderBlockBits);
    if (n == 0) return hipSuccess;
    if (nsmall <= kOrderTrial) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_order_bounds, dim3((kBlocks + 1 + 255) / 256), dim3(256), 0, s, primes, n,
                       kBlocks + 1, bounds);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_order2, dim3(kBlocks), dim3(kOrderThreads), 0, s, primes, bounds, small,
                       nsmall, orders);
    return hipGetLastError();
}

hipError_t launch_primes32(const u32 *small, u32 nsmall, u32 nblocks, u32 *counts,
                           const u64 *offsets, u32 *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_primes32, dim3(nblocks), dim3(kPrimeThreads), 0, s, small, nsmall,
                       counts, offsets, out);
    return hipGetLastError();
}

hipError_t launch_bucket_sieve(const BatchSpan *spans, const u64 *prefix, const BucketSpan *bspans,
                               u32 nb, u32 nbands, const u32 *primes, const u32 *orders,
                               const BucketStore &st, hipStream_t s)
{
    hipLaunchKernelGGL(k_bucket_fill, dim3(kBucketGroups), dim3(kFillThreads), 0, s, spans, prefix,
                       bspans, nb, primes, orders, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || nbands == 0) return e;
    hipLaunchKernelGGL(k_bucket_split, dim3(nbands * kBucketSplitParts), dim3(kFillThreads), 0, s,
                       bspans, nb, st, (u32)kBucketGroups);
    return hipGetLastError();
}

} // namespace ww
