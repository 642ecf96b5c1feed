/*
 * ww_presieve.hip -- the wide pre-sieve of the persistent form.
 *
 * For every 8192-integer segment of a batch it writes 64 words (one per lane of the wave
 * that will sieve the segment, bit b of word w = the odd number base + 128 w + 2 b) of
 * numbers to drop: the odd multiples of the deep primes above the per-segment bound B1 and
 * up to B2 (at most 2^22), except each prime's base-2 pseudoprime class ell mod
 * m = ell * lcm(2, ord_ell(2)), which holds every strong pseudoprime the prime divides.
 * So, as with the per-segment deep sieve, no member of the reference's PRP list is ever
 * dropped; only composites are, before their strong PRP test.
 *
 * Work unit: a chunk of `chunk_segs` consecutive segments (a power of two from 8 to 256:
 * a sieve segment of 2^15 to 2^20 bits, one per odd number; 256 = 2^21 integers by
 * default), split where it crosses from one range of the batch to the next. One workgroup
 * marks a chunk in an LDS bitmap (128 KB at the largest size), so each prime's start
 * positions (two 64-bit Barrett reductions) are computed once per chunk instead of once
 * per 4096 odd numbers. Pseudoprime classes are collected in an LDS list (at most a few
 * hundred per chunk; the table starts above 1024 for that); if the list overflows, the
 * chunk drops nothing (exact, only slower). Pieces shorter than min(16, chunk_segs)
 * segments are not pre-sieved.
 *
 * With the bucket sieve (ww_bucket.hip) the chunk's buckets are applied after those
 * conservative clears, as exact marks, and a second set of words is written: the odd
 * numbers that still need the strong PRP test, i.e. the pseudoprime-class members of every
 * tier, or every odd number of a piece that is not sieved up to sqrt(hi).
 */
#include "ww_arith.h"
#include "ww_batch.h"
#include "ww_bucket.h"

namespace ww {

namespace {

constexpr int kPreThreads = 1024;
constexpr u32 kChunkSegs = kPresieveMaxChunkSegs;
constexpr u32 kChunkPos = kChunkSegs * 4096; /* odd numbers per chunk, at most */
constexpr u32 kChunkWords = kChunkPos / 32;
constexpr u32 kRescueCap = 6144;

/* x mod m for any 64-bit x with bar = floor(2^64 / m): the Barrett quotient is the true
   one or one less. */
__device__ __forceinline__ u64 barrett64(u64 x, u64 m, u64 bar)
{
    u64 r = x - mul_hi(x, bar) * m;
    return r >= m ? r - m : r;
}

/* Largest a with prefix[a] <= s (the span holding global segment s). */
__device__ __forceinline__ u32 span_of(const u64 *prefix, u32 nspans, u64 s)
{
    u32 a = 0, b = nspans;
    while (b - a > 1) {
        const u32 mid = (a + b) >> 1;
        if (prefix[mid] <= s) a = mid;
        else b = mid;
    }
    return a;
}

/* Applies the buckets of chunk j of a bucket span (one per split part) to the LDS bitmap:
   marks the hits and appends the pseudoprime-class hits to the rescue list; sets *over when
   a bucket overflowed (entries were lost). Called by every thread, with 16 loads in flight
   per lane. */
__device__ __forceinline__ void apply_buckets(const BucketArgs &bk, const BucketSpan &bs, u32 j,
                                              u32 *bm, u32 *rescue, u32 *rcount, u32 *over)
{
    constexpr int kUnroll = 16;
    for (u32 p = 0; p < kBucketSplitParts; ++p) {
        const u64 at0 = (u64)j * kBucketSplitParts + p;
        u32 n = bk.counts[bs.cnt_base + at0];
        if (n > bk.cap) {
            if (threadIdx.x == 0) *over = 1;
            n = bk.cap;
        }
        const u32 *run = bk.entries + bs.ent_base + at0 * bk.cap;
        for (u32 i0 = 0; i0 < n; i0 += kPreThreads * kUnroll) {
            u32 e[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                const u32 i = i0 + kPreThreads * u + threadIdx.x;
                e[u] = i < n ? run[i] : 0xffffffffu;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; ++u) {
                if (e[u] == 0xffffffffu) continue;
                const u32 at = e[u] & ((1u << kBucketChunkBits) - 1);
                if (e[u] & kBucketRescue) {
                    const u32 k = atomicAdd(rcount, 1u);
                    if (k < kRescueCap) rescue[k] = at;
                } else {
                    atomicOr(&bm[at >> 5], 1u << (at & 31));
                }
            }
        }
    }
}

__global__ __launch_bounds__(kPreThreads) void k_presieve(const BatchSpan *spans,
                                                          const u64 *prefix, u32 nspans,
                                                          const WidePrime *tab, u32 ntab,
                                                          u64 *out, u32 chunk_segs, u64 *outr,
                                                          BucketArgs bk)
{
    __shared__ __align__(8) u32 bm[kChunkWords]; /* read back as u64 words below */
    __shared__ u32 rescue[kRescueCap];
    __shared__ u32 rcount;
    __shared__ u32 lost; /* the rescue list overflowed with the buckets' class hits */
    __shared__ u32 over; /* a bucket of the current piece overflowed */
    /* segments of the chunk whose candidates all need the strong PRP test (not sieved up
       to sqrt(hi)); only kept when outr is given */
    __shared__ uint8_t all_prp[kChunkSegs];

    const u64 total = prefix[nspans];
    const u64 nchunks = (total + chunk_segs - 1) / chunk_segs;
    const u64 min_piece = chunk_segs < kPresieveMinSegs ? chunk_segs : kPresieveMinSegs;
    const u32 chunk_words = chunk_segs * (4096 / 32);
    for (u64 c = blockIdx.x; c < nchunks; c += gridDim.x) {
        const u64 s0 = c * chunk_segs;
        const u64 s1 = s0 + chunk_segs < total ? s0 + chunk_segs : total;
        for (u32 w = threadIdx.x; w < chunk_words; w += kPreThreads) bm[w] = 0;
        if (threadIdx.x < kChunkSegs) all_prp[threadIdx.x] = 0;
        if (threadIdx.x == 0) {
            rcount = 0;
            lost = 0;
            over = 0;
        }
        __syncthreads();

        /* the chunk's pieces, each inside one range (the same values in every thread) */
        for (u64 s = s0; s < s1;) {
            const u32 a = span_of(prefix, nspans, s);
            const u64 e = prefix[a + 1] < s1 ? prefix[a + 1] : s1;
            const BatchSpan sp = spans[a];
            const u64 G = (sp.lo | 1) + (sp.seg0 + (s - prefix[a])) * kIntsPerWave;
            const u32 off = (u32)(s - s0) * 4096; /* odd-number index of G in the chunk */
            const u32 len = (u32)(e - s) * 4096;
            /* a short piece (a small range, or the end of one) is left to the strong PRP
               test: its words stay zero */
            const bool sieved = e - s >= min_piece;
            const u32 nprimes = sieved ? ntab : 0;
            if (outr && (!sieved || !bk.spans || sp.pad == 0))
                for (u32 k = threadIdx.x; k < (u32)(e - s); k += kPreThreads)
                    all_prp[(u32)(s - s0) + k] = 1;
            /* each prime's entry is loaded one iteration ahead, so the table's latency hides
               behind the marks of the prime before */
            WidePrime dn{};
            if (threadIdx.x < nprimes) dn = tab[threadIdx.x];
            for (u32 j = threadIdx.x; j < nprimes; j += kPreThreads) {
                const WidePrime d = dn;
                if (j + kPreThreads < nprimes) dn = tab[j + kPreThreads];
                const u64 ell = d.ell;
                /* first odd multiple G + 2i of ell: -G mod ell, plus ell when odd */
                const u64 r = barrett64(G, ell, d.bar_ell);
                u64 dd = r ? ell - r : 0;
                if (dd & 1) dd += ell;
                u32 i = (u32)(dd >> 1); /* < ell < 2^22 */
                if (G <= ell && G + 2 * (u64)i == ell) i += (u32)ell; /* never ell itself */
                /* positions in the chunk, below 2^20 + 2^22: 32-bit arithmetic */
                const u32 end = off + len;
                for (u32 at = off + i; at < end; at += (u32)ell)
                    atomicOr(&bm[at >> 5], 1u << (at & 31));
                /* members of the class ell mod m: G + 2k = ell (mod m), both odd */
                const u64 rm = barrett64(G, d.m, d.bar_m);
                for (u64 k = (ell >= rm ? ell - rm : ell + d.m - rm) >> 1; k < len; k += d.m >> 1) {
                    const u32 at = atomicAdd(&rcount, 1u);
                    if (at < kRescueCap) rescue[at] = off + (u32)k;
                }
            }
            s = e;
        }
        __syncthreads();
        const u32 nr = rcount;
        if (nr <= kRescueCap)
            for (u32 k = threadIdx.x; k < nr; k += kPreThreads)
                atomicAnd(&bm[rescue[k] >> 5], ~(1u << (rescue[k] & 31)));
        __syncthreads();
        /* the buckets of the chunk's bucket-span pieces: exact marks (after the conservative
           clears above), their pseudoprime-class hits joining the rescue list */
        if (outr && bk.spans && nr <= kRescueCap) {
            for (u64 s = s0; s < s1;) {
                const u32 a = span_of(prefix, nspans, s);
                const u64 e = prefix[a + 1] < s1 ? prefix[a + 1] : s1;
                const u32 pad = spans[a].pad;
                if (pad && e - s >= min_piece) {
                    const BucketSpan bs = bk.spans[pad - 1];
                    const u32 j = (u32)(s0 >> (kBucketChunkBits - 12)) - bs.chunk_first;
                    apply_buckets(bk, bs, j, bm, rescue, &rcount, &over);
                    __syncthreads();
                    if (over) { /* the same in every thread after the barrier */
                        for (u32 k = threadIdx.x; k < (u32)(e - s); k += kPreThreads)
                            all_prp[(u32)(s - s0) + k] = 1;
                        __syncthreads();
                        if (threadIdx.x == 0) {
                            atomicAdd(bk.overflows, 1u);
                            over = 0;
                        }
                    }
                    __syncthreads();
                }
                s = e;
            }
            __syncthreads();
            if (threadIdx.x == 0 && rcount > kRescueCap) lost = 1;
        }
        __syncthreads();
        const u64 *bm64 = reinterpret_cast<const u64 *>(bm);
        const u32 nw = (u32)(s1 - s0) * 64;
        for (u32 w = threadIdx.x; w < nw; w += kPreThreads)
            out[s0 * 64 + w] = nr <= kRescueCap ? bm64[w] : 0;
        if (outr) {
            /* the words of candidates that still need the strong PRP test: the rescue list
               (pseudoprime classes of every tier), or everything where the chunk is not
               fully sieved */
            const bool all = nr > kRescueCap || lost;
            const u32 nrr = rcount < kRescueCap ? rcount : kRescueCap;
            __syncthreads();
            for (u32 w = threadIdx.x; w < chunk_words; w += kPreThreads) bm[w] = 0;
            __syncthreads();
            for (u32 k = threadIdx.x; k < nrr; k += kPreThreads)
                atomicOr(&bm[rescue[k] >> 5], 1u << (rescue[k] & 31));
            __syncthreads();
            for (u32 w = threadIdx.x; w < nw; w += kPreThreads)
                outr[s0 * 64 + w] = (all || all_prp[w >> 6]) ? ~0ull : bm64[w];
        }
        __syncthreads();
    }
}

/* The single span [lo, hi) of a plain-form chunk, written on the device so that the
   pre-sieve of a chunk needs no host copy. */
__global__ void k_set_span(BatchSpan *span, u64 *prefix, u64 lo, u64 hi)
{
    span->lo = lo;
    span->hi = hi;
    span->seg0 = 0;
    span->slot = 0;
    span->pad = 0;
    span->pre_seg = 0;
    prefix[0] = 0;
    prefix[1] = segments_of(lo, hi);
}

/* The same for a fused chunk with the bucket sieve: the span refers to bucket span 0
   (pad = 1), which is written beside it (chunk 0 at the span's first odd number). */
__global__ void k_set_bucket_span(BatchSpan *span, u64 *prefix, BucketSpan *bspan, u64 lo,
                                  u64 hi, BucketSpan bs)
{
    span->lo = lo;
    span->hi = hi;
    span->seg0 = 0;
    span->slot = 0;
    span->pad = 1;
    span->pre_seg = 0;
    prefix[0] = 0;
    prefix[1] = segments_of(lo, hi);
    *bspan = bs;
}

} // namespace

hipError_t launch_set_span(u64 lo, u64 hi, BatchSpan *span, u64 *prefix, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_span, dim3(1), dim3(1), 0, s, span, prefix, lo, hi);
    return hipGetLastError();
}

hipError_t launch_set_bucket_span(u64 lo, u64 hi, BatchSpan *span, u64 *prefix, BucketSpan *bspan,
                                  const BucketSpan &bs, hipStream_t s)
{
    hipLaunchKernelGGL(k_set_bucket_span, dim3(1), dim3(1), 0, s, span, prefix, bspan, lo, hi, bs);
    return hipGetLastError();
}

hipError_t launch_presieve_range(u64 lo, u64 hi, BatchSpan *span, u64 *prefix,
                                 const WidePrime *tab, u32 ntab, u64 *out, int grid,
                                 hipStream_t s, u32 chunk_segs)
{
    hipError_t e = launch_set_span(lo, hi, span, prefix, s);
    if (e != hipSuccess) return e;
    return launch_presieve(span, prefix, 1, tab, ntab, out, grid, s, chunk_segs);
}

bool presieve_chunk_ok(u32 chunk_segs)
{
    return chunk_segs >= kPresieveMinChunkSegs && chunk_segs <= kPresieveMaxChunkSegs &&
           (chunk_segs & (chunk_segs - 1)) == 0;
}

hipError_t launch_presieve(const BatchSpan *spans, const u64 *prefix, u32 nspans,
                           const WidePrime *tab, u32 ntab, u64 *out, int grid, hipStream_t s,
                           u32 chunk_segs, u64 *outr, const BucketArgs *bk)
{
    if (!presieve_chunk_ok(chunk_segs)) return hipErrorInvalidValue;
    BucketArgs b{};
    if (bk && bk->spans) {
        if (!outr || chunk_segs != (1u << (kBucketChunkBits - 12)) || !bk->overflows)
            return hipErrorInvalidValue;
        b = *bk;
    }
    hipLaunchKernelGGL(k_presieve, dim3(grid), dim3(kPreThreads), 0, s, spans, prefix, nspans,
                       tab, ntab, out, chunk_segs, outr, b);
    return hipGetLastError();
}

} // namespace ww
