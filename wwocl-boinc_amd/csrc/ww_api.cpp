/*
 * ww_api.cpp -- the C ABI of include/ww_api.h over HIP.
 *
 * One context = one device, one stream, the device counters, the PRP list and the
 * result records (the cl_mem set of progData, main.cpp:51-66). Every HIP error is
 * turned into a negative return code and a thread-local message; nothing here falls
 * back to the CPU: without a usable GPU every device entry point fails.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/ww_api.h"
#include "ww_arith.h"
#include "ww_batch.h"
#include "ww_bucket.h"
static_assert(sizeof(ww_deep_prime) == sizeof(ww::DeepPrime), "deep table layout");
#include "ww_kernels.h"

using ww::BatchSpan;
using ww::BatchState;
using ww::DevState;
using ww::SlotSum;

struct ww_ctx {
    /* held by every entry point that takes the context (recursive: the autotuner and the
       range loop call other entry points), so calls on one context from several host
       threads are serialised */
    std::recursive_mutex mu;
    int device = 0;
    hipStream_t stream = nullptr;
    DevState *d_state = nullptr;
    uint64_t *d_list = nullptr;
    uint64_t list_cap = 0;
    ww_result *d_results = nullptr;
    uint32_t result_cap = 0;
    int test_grid = 0;
    int cus = 0;
    ww::DeepPrime *d_deep = nullptr; /* primes 127..kMaxDeepBound, ascending */
    uint32_t ndeep_max = 0;
    uint32_t ndeep = 0; /* entries in use: primes <= the deep-sieve bound */
    uint32_t deep_bound = 0;
    std::vector<uint32_t> deep_ells;
    /* wide pre-sieve of the persistent form: primes 127..WW_MAX_PRESIEVE_BOUND, built on
       first use; the active ones are deep_bound < ell <= presieve_bound */
    ww::WidePrime *d_wide = nullptr;
    std::vector<uint32_t> wide_ells;
    uint32_t wide_built = 0;   /* the table holds the primes up to this bound */
    uint32_t presieve_bound = 0;
    /* plain form: pre-sieve words of the current chunk, and its span on the device */
    uint64_t *d_pre_plain = nullptr;
    uint64_t pre_plain_segs = 0;
    ww::BatchSpan *d_span1 = nullptr;
    uint64_t *d_prefix1 = nullptr;
    /* fused chunks (ww_search_fused): the persistent kernel's state, sum and records */
    ww::BatchState *d_fstate = nullptr;
    SlotSum *d_fsum = nullptr;
    ww_result *d_fres = nullptr;
    /* single chunks with the bucket sieve (ww_search_fused, ww_getsegprimes): band and
       chunk buckets for up to fb_segs segments
       (sized for sieving primes up to 2^32 at capacity scale fb_scale), the bucket span,
       the overflow count and the needs-PRP words */
    ww::BucketSpan *d_fbspan = nullptr;
    uint32_t *d_fband_ent = nullptr, *d_fband_cnt = nullptr;
    uint32_t *d_fbent = nullptr, *d_fbcnt = nullptr, *d_fbover = nullptr;
    uint64_t *d_fprer = nullptr;
    uint32_t fband_cap = 0, fbcap = 0;
    uint64_t fb_segs = 0;
    double fb_scale = 0;
    int tuned_blocks = 0;      /* defaults for batches created from now on */
    uint32_t tuned_claim = 8;
    /* launch shape for launches issued from now on: waves per workgroup of the persistent
       kernel, and the pre-sieve's segment (in 8192-integer segments) */
    int waves = ww::kDefaultBatchWaves;
    uint32_t presieve_chunk = ww::kPresieveMaxChunkSegs;
    ww_tuning last_tuning{}; /* the tuner's (or a loaded cache's) best, for ww_tuning_save */
    uint64_t default_chunk = 0;
    /* bucket sieve: every odd prime below 2^32 and the base-2 order of each one above 2^20
       (built on first use), the number of table primes in each 2^21-integer block before
       it (for index bounds), and the smallest range width that gets the bucket sieve in
       batches created from now on */
    uint32_t *d_primes32 = nullptr;
    uint32_t *d_orders32 = nullptr;
    uint64_t nprimes32 = 0;
    std::vector<uint64_t> primes32_prefix;
    uint64_t bucket_min_width = WW_DEFAULT_BUCKET_MIN_WIDTH;
    double bucket_scale = 1.0;
    /* event pairs of launches whose durations have not been collected yet */
    std::vector<hipEvent_t> ev_sieve, ev_tests;
    double ms_sieve = 0, ms_tests = 0;
    uint64_t launches = 0;
};

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (expr);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(WW_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),  \
                        __FILE__, __LINE__);                                             \
    } while (0)

/* Holds the context's lock for the rest of the calling entry point. */
struct CtxLock {
    std::unique_lock<std::recursive_mutex> l;
    explicit CtxLock(ww_ctx *c)
    {
        if (c) l = std::unique_lock<std::recursive_mutex>(c->mu);
    }
};
#define WW_LOCK(c) CtxLock ctx_lock_(c)

int use(ww_ctx *c)
{
    if (!c) return fail(WW_E_ARG, "null context");
    HIP_TRY(hipSetDevice(c->device));
    return WW_OK;
}

/* Record a start/stop event pair around one launch for ww_kernel_times. */
int event_pair(ww_ctx *c, std::vector<hipEvent_t> &v, hipEvent_t *a, hipEvent_t *b)
{
    HIP_TRY(hipEventCreate(a));
    HIP_TRY(hipEventCreate(b));
    v.push_back(*a);
    v.push_back(*b);
    return WW_OK;
}

int collect(std::vector<hipEvent_t> &v, double *acc)
{
    for (size_t i = 0; i + 1 < v.size(); i += 2) {
        float ms = 0;
        HIP_TRY(hipEventSynchronize(v[i + 1]));
        HIP_TRY(hipEventElapsedTime(&ms, v[i], v[i + 1]));
        *acc += ms;
        (void)hipEventDestroy(v[i]);
        (void)hipEventDestroy(v[i + 1]);
    }
    v.clear();
    return WW_OK;
}

/* Pending event pairs are collected once this many accumulate, so a long run holds a
   bounded number of events. */
constexpr size_t kMaxPendingEvents = 512;

/* Active pre-sieve primes for this context: [first, first + count) of the wide table
   (deep_bound < ell <= presieve_bound), as one number for graph bookkeeping. */
void wide_span(ww_ctx *c, uint32_t *first, uint32_t *count)
{
    *first = *count = 0;
    if (!c->d_wide || !c->presieve_bound) return;
    /* above the per-segment deep sieve, and above 1024 in any case: smaller primes have
       small pseudoprime moduli m, whose classes would overflow the pre-sieve's list */
    uint32_t lo = c->ndeep ? c->deep_bound : 0;
    if (lo < ww::kPresieveMinPrime) lo = ww::kPresieveMinPrime;
    uint32_t a = 0, e = 0;
    while (a < c->wide_ells.size() && c->wide_ells[a] <= lo) ++a;
    e = a;
    while (e < c->wide_ells.size() && c->wide_ells[e] <= c->presieve_bound) ++e;
    *first = a;
    *count = e - a;
}

/* Builds the table of every odd prime below 2^32 on the device (once per context): a
   counting pass of k_primes32, the exclusive prefix of the per-block counts on the host,
   then the writing pass. */
int ensure_primes32(ww_ctx *c)
{
    if (c->d_primes32) return WW_OK;
    constexpr uint32_t kBlocks = 2048; /* 2^21 integers each */
    std::vector<uint32_t> small;
    {
        std::vector<uint8_t> comp(1u << 16, 0);
        for (uint32_t i = 3; i < (1u << 16); i += 2) {
            if (comp[i]) continue;
            small.push_back(i);
            for (uint32_t j = i * i; j < (1u << 16); j += 2 * i) comp[j] = 1;
        }
    }
    uint32_t *d_small = nullptr, *d_counts = nullptr;
    uint64_t *d_off = nullptr;
    std::vector<uint32_t> counts(kBlocks);
    std::vector<uint64_t> off(kBlocks + 1, 0);
    hipError_t e = hipMalloc(&d_small, small.size() * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&d_counts, kBlocks * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&d_off, kBlocks * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_small, small.data(), small.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = ww::launch_primes32(d_small, (uint32_t)small.size(), kBlocks, d_counts, nullptr, nullptr,
                                c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(counts.data(), d_counts, kBlocks * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) {
        for (uint32_t b = 0; b < kBlocks; ++b) off[b + 1] = off[b] + counts[b];
        e = hipMalloc(&c->d_primes32, off[kBlocks] * sizeof(uint32_t));
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_off, off.data(), kBlocks * sizeof(uint64_t), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess)
        e = ww::launch_primes32(d_small, (uint32_t)small.size(), kBlocks, d_counts, d_off,
                                c->d_primes32, c->stream);
    /* the base-2 orders of the primes above 2^20 */
    ww::SmallDiv *d_div = nullptr;
    uint64_t *d_bounds = nullptr;
    if (e == hipSuccess) e = hipMalloc(&c->d_orders32, off[kBlocks] * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&d_div, ww::kSmallOddPrimes * sizeof(ww::SmallDiv));
    if (e == hipSuccess) e = hipMalloc(&d_bounds, ((1u << 16) + 1) * sizeof(uint64_t));
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_div, ww::small_div_table(), ww::kSmallOddPrimes * sizeof(ww::SmallDiv),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = ww::launch_order2(c->d_primes32, off[kBlocks], d_div, ww::kSmallOddPrimes, c->d_orders32,
                              d_bounds, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (d_small) (void)hipFree(d_small);
    if (d_counts) (void)hipFree(d_counts);
    if (d_off) (void)hipFree(d_off);
    if (d_div) (void)hipFree(d_div);
    if (d_bounds) (void)hipFree(d_bounds);
    if (e != hipSuccess) {
        if (c->d_primes32) (void)hipFree(c->d_primes32);
        if (c->d_orders32) (void)hipFree(c->d_orders32);
        c->d_primes32 = nullptr;
        c->d_orders32 = nullptr;
        return fail(WW_E_HIP, "prime table: %s", hipGetErrorString(e));
    }
    c->nprimes32 = off[kBlocks];
    c->primes32_prefix = off;
    return WW_OK;
}

/* isqrt(x) for 64-bit x. */
uint64_t isqrt64(uint64_t x)
{
    uint64_t r = (uint64_t)std::sqrt((double)x);
    while (r > 0 && (r > 0xffffffffull || r * r > x)) --r;
    while (r < 0xffffffffull && (r + 1) * (r + 1) <= x) ++r;
    return r;
}

/* Whether the context's sieve tiers allow known primes (no strong PRP test for candidates
   sieved up to sqrt(hi)): the deep sieve reaches the pre-sieve's lowest prime (1024), the
   pre-sieve reaches 2^20 (the bucket sieve's lowest prime, which its capacities assume) and
   marks 2^20-bit chunks (the buckets' unit). */
bool bucket_config_ok(const ww_ctx *c)
{
    return c->ndeep && c->deep_bound >= ww::kPresieveMinPrime &&
           c->presieve_bound >= (1u << ww::kBucketChunkBits) &&
           c->presieve_chunk == (1u << (ww::kBucketChunkBits - 12));
}

/* Table index of the first prime above the pre-sieve bound: the odd primes up to 113 (29)
   and the wide table's primes up to the bound. */
/* The share of a sieving prime's odd multiples that the bucket walk visits: cofactors
   prime to 30030 (5760 of 15015 odd ones), and to the primes from 17 to 53. */
constexpr double kVisited = 5760.0 / 15015.0 * 16.0 / 17.0 * 18.0 / 19.0 * 22.0 / 23.0 *
                            28.0 / 29.0 * 30.0 / 31.0 * 36.0 / 37.0 * 40.0 / 41.0 *
                            42.0 / 43.0 * 46.0 / 47.0 * 52.0 / 53.0;

uint64_t bucket_first_index(const ww_ctx *c)
{
    uint64_t n = 29;
    for (uint32_t l : c->wide_ells) {
        if (l > c->presieve_bound) break;
        ++n;
    }
    return n;
}

} // namespace

namespace {

/* The device span of a single chunk, and its pre-sieve words for nseg segments (grown
   when needed; queued launches may still read the old buffer, so the stream is drained
   first). */
int ensure_chunk_span(ww_ctx *c)
{
    if (c->d_span1) return WW_OK;
    hipError_t e;
    if ((e = hipMalloc(&c->d_prefix1, 2 * sizeof(uint64_t))) != hipSuccess ||
        (e = hipMalloc(&c->d_span1, sizeof(ww::BatchSpan))) != hipSuccess) { /* span last */
        if (c->d_prefix1) (void)hipFree(c->d_prefix1);
        c->d_prefix1 = nullptr;
        return fail(WW_E_HIP, "chunk span: %s", hipGetErrorString(e));
    }
    return WW_OK;
}

int ensure_pre_plain(ww_ctx *c, uint64_t nseg)
{
    if (nseg <= c->pre_plain_segs) return WW_OK;
    hipError_t e;
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return fail(WW_E_HIP, "pre-sieve buffer: %s", hipGetErrorString(e));
    if (c->d_pre_plain) (void)hipFree(c->d_pre_plain);
    c->d_pre_plain = nullptr;
    c->pre_plain_segs = 0;
    if ((e = hipMalloc(&c->d_pre_plain, nseg * ww::kPresieveWordsPerSeg * sizeof(uint64_t))) != hipSuccess)
        return fail(WW_E_HIP, "pre-sieve buffer: %s", hipGetErrorString(e));
    c->pre_plain_segs = nseg;
    return WW_OK;
}

/* The one-span table of a single chunk [low, high) on the device, and its wide pre-sieve
   when the pre-sieve is on and the chunk is long enough: *pre = the words, else null. */
int chunk_presieve(ww_ctx *c, uint64_t low, uint64_t high, const uint64_t **pre)
{
    *pre = nullptr;
    hipError_t e;
    int rc = ensure_chunk_span(c);
    if (rc) return rc;
    uint32_t wfirst, wcount;
    wide_span(c, &wfirst, &wcount);
    const uint64_t nseg = ww::segments_of(low, high);
    if (!wcount || nseg < ww::kPresieveMinSegs) {
        if ((e = ww::launch_set_span(low, high, c->d_span1, c->d_prefix1, c->stream)) != hipSuccess)
            return fail(WW_E_HIP, "chunk span: %s", hipGetErrorString(e));
        return WW_OK;
    }
    if ((rc = ensure_pre_plain(c, nseg))) return rc;
    const int pgrid = (int)std::min<uint64_t>((nseg + c->presieve_chunk - 1) / c->presieve_chunk,
                                              (uint64_t)c->cus);
    if ((e = ww::launch_presieve_range(low, high, c->d_span1, c->d_prefix1, c->d_wide + wfirst, wcount,
                                       c->d_pre_plain, pgrid > 0 ? pgrid : 1, c->stream,
                                       c->presieve_chunk)) != hipSuccess)
        return fail(WW_E_HIP, "pre-sieve: %s", hipGetErrorString(e));
    *pre = c->d_pre_plain;
    return WW_OK;
}

void free_fused_buckets(ww_ctx *c)
{
    for (uint32_t **p : {&c->d_fband_ent, &c->d_fband_cnt, &c->d_fbent, &c->d_fbcnt, &c->d_fbover}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
    }
    if (c->d_fprer) (void)hipFree(c->d_fprer);
    if (c->d_fbspan) (void)hipFree(c->d_fbspan);
    c->d_fprer = nullptr;
    c->d_fbspan = nullptr;
    c->fb_segs = 0;
}

/* The bucket storage of fused chunks of up to nseg segments, sized as a batch's for a top
   of 2^32 (so any chunk of that length fits) and the context's capacity scale. When the
   device memory is not there, *avail = false and the chunk runs without it (the failed
   allocation's error is cleared, so later launch checks do not see it). */
int ensure_fused_buckets(ww_ctx *c, uint64_t nseg, bool *avail)
{
    *avail = true;
    if (nseg <= c->fb_segs && c->fb_scale == c->bucket_scale) return WW_OK;
    hipError_t e = hipStreamSynchronize(c->stream); /* queued chunks may use the old ones */
    if (e != hipSuccess) return fail(WW_E_HIP, "fused buckets: %s", hipGetErrorString(e));
    free_fused_buckets(c);
    const uint64_t chunks = ((nseg * 4096 - 1) >> ww::kBucketChunkBits) + 1;
    const uint64_t bands = (chunks + ww::kBucketBandChunks - 1) / ww::kBucketBandChunks;
    const double mean = std::ldexp(1.0, 20) * kVisited * std::log(32.0 / 20.0);
    const double sc = c->bucket_scale;
    c->fbcap = (uint32_t)std::ceil(sc * (1.15 * mean / ww::kBucketSplitParts + 512));
    c->fband_cap = (uint32_t)std::ceil(sc * (1.25 * mean * ww::kBucketBandChunks / ww::kBucketGroups + 256));
    const size_t band_buckets = bands * ww::kBucketGroups, parts = chunks * ww::kBucketSplitParts;
    e = hipMalloc(&c->d_fband_cnt, band_buckets * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_fband_ent, band_buckets * (size_t)c->fband_cap * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_fbcnt, parts * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_fbent, parts * (size_t)c->fbcap * sizeof(uint32_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_fbover, sizeof(uint32_t));
    if (e == hipSuccess) e = hipMemsetAsync(c->d_fbover, 0, sizeof(uint32_t), c->stream);
    if (e == hipSuccess) e = hipMalloc(&c->d_fprer, nseg * ww::kPresieveWordsPerSeg * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&c->d_fbspan, sizeof(ww::BucketSpan)); /* last: all made */
    if (e == hipErrorOutOfMemory) {
        (void)hipGetLastError();
        free_fused_buckets(c);
        *avail = false;
        return WW_OK;
    }
    if (e != hipSuccess) {
        free_fused_buckets(c);
        return fail(WW_E_HIP, "fused buckets: %s", hipGetErrorString(e));
    }
    c->fb_segs = nseg;
    c->fb_scale = sc;
    return WW_OK;
}

/* A single chunk (of ww_search_fused or ww_getsegprimes) with the bucket sieve, when
   [low, high) is at least the bucket width, the sieve tiers allow known primes and the chunk
   is pre-sieved: its span and bucket span, the fill and split, and the pre-sieve with the
   needs-PRP words; *pre and *prer = the two sets of words. Both stay null when the chunk
   does not qualify (the plain pre-sieve then runs). */
int chunk_buckets(ww_ctx *c, uint64_t low, uint64_t high, const uint64_t **pre,
                  const uint64_t **prer)
{
    *pre = *prer = nullptr;
    uint32_t wfirst, wcount;
    wide_span(c, &wfirst, &wcount);
    const uint64_t nseg = ww::segments_of(low, high);
    if (!wcount || nseg < ww::kPresieveMinSegs || high - low < c->bucket_min_width || !bucket_config_ok(c))
        return WW_OK;
    const uint64_t chunks = ((nseg * 4096 - 1) >> ww::kBucketChunkBits) + 1;
    if (chunks > ww::kBucketMaxChunks) return WW_OK;
    int rc;
    bool avail = false;
    if ((rc = ensure_primes32(c)) || (rc = ensure_chunk_span(c)) || (rc = ensure_pre_plain(c, nseg)) ||
        (rc = ensure_fused_buckets(c, nseg, &avail)))
        return rc;
    if (!avail) return WW_OK;
    ww::BucketSpan bs{};
    bs.span = 0;
    bs.bit0 = 0;
    bs.chunk_first = 0;
    bs.nchunks = (uint32_t)chunks;
    bs.nbands = (uint32_t)((chunks + ww::kBucketBandChunks - 1) / ww::kBucketBandChunks);
    bs.top = (uint32_t)isqrt64(high - 1);
    bs.pidx0 = bucket_first_index(c);
    bs.pidx1 = c->primes32_prefix[std::min<uint64_t>(((uint64_t)bs.top >> 21) + 1,
                                                     c->primes32_prefix.size() - 1)];
    if (bs.pidx1 < bs.pidx0) bs.pidx1 = bs.pidx0;
    HIP_TRY(ww::launch_set_bucket_span(low, high, c->d_span1, c->d_prefix1, c->d_fbspan, bs, c->stream));
    const ww::BucketStore st{c->d_fband_ent, c->d_fband_cnt, c->fband_cap, c->d_fbent, c->d_fbcnt, c->fbcap};
    HIP_TRY(ww::launch_bucket_sieve(c->d_span1, c->d_prefix1, c->d_fbspan, 1, bs.nbands, c->d_primes32,
                                    c->d_orders32, st, c->stream));
    const ww::BucketArgs bk{c->d_fbspan, c->d_fbent, c->d_fbcnt, c->fbcap, c->d_fbover};
    const int pgrid = (int)std::min<uint64_t>((nseg + c->presieve_chunk - 1) / c->presieve_chunk,
                                              (uint64_t)c->cus);
    HIP_TRY(ww::launch_presieve(c->d_span1, c->d_prefix1, 1, c->d_wide + wfirst, wcount, c->d_pre_plain,
                                pgrid > 0 ? pgrid : 1, c->stream, c->presieve_chunk, c->d_fprer, &bk));
    *pre = c->d_pre_plain;
    *prer = c->d_fprer;
    return WW_OK;
}

} // namespace

int ww_set_error(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return code;
}

/* Context internals the serving engine needs (same library). */
int ww_ctx_internals(ww_ctx *c, int *device, const ww::DeepPrime **deep, uint32_t *ndeep,
                     int *cus, int *tuned_blocks, uint32_t *tuned_claim,
                     const ww::WidePrime **wide, uint32_t *nwide, int *waves,
                     uint32_t *presieve_chunk)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    uint32_t wfirst, wcount;
    wide_span(c, &wfirst, &wcount);
    *wide = wcount ? c->d_wide + wfirst : nullptr;
    *nwide = wcount;
    *device = c->device;
    *deep = c->d_deep;
    *ndeep = c->ndeep;
    *cus = c->cus;
    *tuned_blocks = c->tuned_blocks;
    *tuned_claim = c->tuned_claim;
    *waves = c->waves;
    *presieve_chunk = c->presieve_chunk;
    return WW_OK;
}

extern "C" {

const char *ww_last_error(void) { return g_err; }

int ww_device_count(int *count)
{
    if (!count) return fail(WW_E_ARG, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *count = n;
    return WW_OK;
}

int ww_device_info(int device, ww_device_info_t *info)
{
    if (!info) return fail(WW_E_ARG, "null info");
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, device));
    memset(info, 0, sizeof *info);
    snprintf(info->name, sizeof info->name, "%s", p.name);
    if (!info->name[0]) (void)hipDeviceGetName(info->name, (int)sizeof info->name, device);
    snprintf(info->arch, sizeof info->arch, "%s", p.gcnArchName);
    if (!info->name[0]) { /* some runtimes report no marketing name: use the target */
        const char *colon = strchr(info->arch, ':');
        const int n = colon ? (int)(colon - info->arch) : (int)strlen(info->arch);
        snprintf(info->name, sizeof info->name, "AMD GPU %.*s", n, info->arch);
    }
    info->compute_units = p.multiProcessorCount;
    info->clock_khz = p.clockRate;
    info->global_mem = (uint64_t)p.totalGlobalMem;
    info->warp_size = p.warpSize;
    return WW_OK;
}

uint64_t ww_list_bound(uint64_t low, uint64_t high)
{
    if (high <= low) return 0;
    double w = (double)(high - low);
    /* Montgomery-Vaughan: pi(x+w) - pi(x) <= 2w/ln w for w > 1 (proven); small windows
       are bounded by the count of odd numbers instead. The base-2 strong pseudoprimes on
       top (about 3.1e7 below 2^64 in all) get the slack x1.02 + 4096: a heuristic, not a
       bound. An overflow past it still shows in the counters (primecount above the
       capacity), which the drop-in host checks after every chunk. */
    double primes = w < 64 ? w : 2.0 * w / std::log(w);
    double odd = w / 2 + 1;
    double b = std::fmin(primes, odd) * 1.02 + 4096;
    return (uint64_t)b;
}

int ww_open(int device, uint64_t list_capacity, uint32_t result_capacity, ww_ctx **out)
{
    if (!out) return fail(WW_E_ARG, "null out");
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0)
        return fail(WW_E_NODEV, "no GPU visible to HIP");
    if (device < 0 || device >= n) return fail(WW_E_ARG, "device %d out of range", device);
    if (list_capacity == 0 || result_capacity == 0)
        return fail(WW_E_ARG, "capacities must be nonzero");
    ww_ctx *c = new ww_ctx;
    c->device = device;
    c->list_cap = list_capacity;
    c->result_cap = result_capacity;
    auto bail = [&](int rc) {
        ww_close(c);
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess) return bail(fail(WW_E_HIP, "hipSetDevice"));
    {
        /* the allocation-free entry point's scratch address on this device */
        const hipError_t ne = ww::noalloc_init(device);
        if (ne != hipSuccess)
            return bail(fail(WW_E_HIP, "scratch address (hipGetSymbolAddress): %s", hipGetErrorString(ne)));
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipGetDeviceProperties"));
    c->cus = prop.multiProcessorCount;
    {
        int nb = ww::tests_blocks_per_cu();
        c->test_grid = prop.multiProcessorCount * (nb > 0 ? nb : 4);
    }
    c->default_chunk = 1ull << 32;
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e)));
    if ((e = hipMalloc(&c->d_state, sizeof(DevState))) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipMalloc(state): %s", hipGetErrorString(e)));
    if ((e = hipMalloc(&c->d_list, list_capacity * sizeof(uint64_t))) != hipSuccess)
        return bail(fail(WW_E_CAPACITY, "hipMalloc(list of %llu): %s",
                         (unsigned long long)list_capacity, hipGetErrorString(e)));
    if ((e = hipMalloc(&c->d_results, (size_t)result_capacity * sizeof(ww_result))) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipMalloc(results): %s", hipGetErrorString(e)));
    {
        std::vector<ww::DeepPrime> t(ww::kMaxDeepBound / 8);
        t.resize(ww_deep_table(ww::kMaxDeepBound, reinterpret_cast<ww_deep_prime *>(t.data()),
                               (uint32_t)t.size()));
        c->ndeep_max = (uint32_t)t.size();
        for (const ww::DeepPrime &d : t) c->deep_ells.push_back(d.ell);
        if ((e = hipMalloc(&c->d_deep, t.size() * sizeof(ww::DeepPrime))) != hipSuccess)
            return bail(fail(WW_E_HIP, "hipMalloc(deep): %s", hipGetErrorString(e)));
        if ((e = hipMemcpyAsync(c->d_deep, t.data(), t.size() * sizeof(ww::DeepPrime),
                                hipMemcpyHostToDevice, c->stream)) != hipSuccess)
            return bail(fail(WW_E_HIP, "hipMemcpy(deep): %s", hipGetErrorString(e)));
        ww_set_deep_bound(c, WW_DEFAULT_DEEP_BOUND);
    }
    if ((e = hipMemsetAsync(c->d_state, 0, sizeof(DevState), c->stream)) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipMemsetAsync: %s", hipGetErrorString(e)));
    {
        const int prc = ww_set_presieve_bound(c, WW_DEFAULT_PRESIEVE_BOUND);
        if (prc) return bail(prc);
    }
    if ((e = hipStreamSynchronize(c->stream)) != hipSuccess)
        return bail(fail(WW_E_HIP, "hipStreamSynchronize: %s", hipGetErrorString(e)));
    *out = c;
    return WW_OK;
}

int ww_close(ww_ctx *c)
{
    if (!c) return WW_OK;
    {
    WW_LOCK(c);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipEvent_t e : c->ev_sieve) (void)hipEventDestroy(e);
    for (hipEvent_t e : c->ev_tests) (void)hipEventDestroy(e);
    if (c->d_deep) (void)hipFree(c->d_deep);
    if (c->d_wide) (void)hipFree(c->d_wide);
    if (c->d_primes32) (void)hipFree(c->d_primes32);
    if (c->d_orders32) (void)hipFree(c->d_orders32);
    if (c->d_fstate) (void)hipFree(c->d_fstate);
    if (c->d_fsum) (void)hipFree(c->d_fsum);
    if (c->d_fres) (void)hipFree(c->d_fres);
    if (c->d_pre_plain) (void)hipFree(c->d_pre_plain);
    free_fused_buckets(c);
    if (c->d_span1) (void)hipFree(c->d_span1);
    if (c->d_prefix1) (void)hipFree(c->d_prefix1);
    if (c->d_results) (void)hipFree(c->d_results);
    if (c->d_list) (void)hipFree(c->d_list);
    if (c->d_state) (void)hipFree(c->d_state);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return WW_OK;
}

int ww_set_deep_bound(ww_ctx *c, uint32_t bound)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    if (bound > ww::kMaxDeepBound) return fail(WW_E_ARG, "deep bound above %u", ww::kMaxDeepBound);
    uint32_t n = 0;
    while (n < c->deep_ells.size() && c->deep_ells[n] <= bound) ++n;
    c->ndeep = n;
    c->deep_bound = bound;
    return WW_OK;
}

int ww_get_deep_bound(ww_ctx *c, uint32_t *bound)
{
    WW_LOCK(c);
    if (!c || !bound) return fail(WW_E_ARG, "null argument");
    *bound = c->deep_bound;
    return WW_OK;
}

int ww_set_presieve_bound(ww_ctx *c, uint32_t bound)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (bound > WW_MAX_PRESIEVE_BOUND)
        return fail(WW_E_ARG, "pre-sieve bound above %u", WW_MAX_PRESIEVE_BOUND);
    if (bound > c->wide_built) {
        /* (re)build the table up to this bound; queued launches may still read the old one */
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->d_wide) (void)hipFree(c->d_wide);
        c->d_wide = nullptr;
        c->wide_ells.clear();
        c->wide_built = 0;
        std::vector<ww_wide_prime> t(bound / 8 + 64);
        t.resize(ww_presieve_table(113, bound, t.data(), (uint32_t)t.size()));
        static_assert(sizeof(ww_wide_prime) == sizeof(ww::WidePrime), "layout");
        hipError_t e = hipMalloc(&c->d_wide, t.size() * sizeof(ww::WidePrime));
        if (e == hipSuccess)
            e = hipMemcpy(c->d_wide, t.data(), t.size() * sizeof(ww::WidePrime), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            if (c->d_wide) (void)hipFree(c->d_wide);
            c->d_wide = nullptr;
            c->presieve_bound = 0; /* no table: the pre-sieve is off */
            return fail(WW_E_HIP, "pre-sieve table: %s", hipGetErrorString(e));
        }
        for (const ww_wide_prime &d : t) c->wide_ells.push_back(d.ell);
        c->wide_built = bound;
    }
    c->presieve_bound = bound;
    return WW_OK;
}

int ww_get_presieve_bound(ww_ctx *c, uint32_t *bound)
{
    WW_LOCK(c);
    if (!c || !bound) return fail(WW_E_ARG, "null argument");
    *bound = c->presieve_bound;
    return WW_OK;
}

int ww_set_bucket_sieve(ww_ctx *c, uint64_t min_width)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    c->bucket_min_width = min_width;
    return WW_OK;
}

int ww_get_bucket_sieve(ww_ctx *c, uint64_t *min_width)
{
    WW_LOCK(c);
    if (!c || !min_width) return fail(WW_E_ARG, "null argument");
    *min_width = c->bucket_min_width;
    return WW_OK;
}

int ww_set_bucket_capacity(ww_ctx *c, double scale)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    if (!(scale > 0 && scale <= 16)) return fail(WW_E_ARG, "bucket capacity scale %g outside (0, 16]", scale);
    c->bucket_scale = scale;
    return WW_OK;
}

/* Entries [first, first + n) of the prime table (orders = false) or of the order table. */
static int read_prime_table(ww_ctx *c, bool orders, uint64_t first, uint64_t n, uint32_t *out,
                            uint64_t *total)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if ((n && !out) || !total) return fail(WW_E_ARG, "null argument");
    if ((rc = ensure_primes32(c))) return rc;
    *total = c->nprimes32;
    if (first > c->nprimes32 || n > c->nprimes32 - first)
        return fail(WW_E_ARG, "entries [%llu, +%llu) beyond the table's %llu", (unsigned long long)first,
                    (unsigned long long)n, (unsigned long long)c->nprimes32);
    if (n) {
        const uint32_t *src = (orders ? c->d_orders32 : c->d_primes32) + first;
        HIP_TRY(hipMemcpyAsync(out, src, n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return WW_OK;
}

int ww_prime_table(ww_ctx *c, uint64_t first, uint64_t n, uint32_t *out, uint64_t *total)
{
    return read_prime_table(c, false, first, n, out, total);
}

int ww_prime_orders(ww_ctx *c, uint64_t first, uint64_t n, uint32_t *out, uint64_t *total)
{
    return read_prime_table(c, true, first, n, out, total);
}

int ww_set_launch_shape(ww_ctx *c, uint32_t waves_per_group, uint32_t segment_bits)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    if (!ww::batch_waves_ok((int)waves_per_group))
        return fail(WW_E_ARG, "waves per workgroup %u: 2, 4 or 8", waves_per_group);
    if (segment_bits < 15 || segment_bits > 20)
        return fail(WW_E_ARG, "segment bits %u outside 15..20", segment_bits);
    c->waves = (int)waves_per_group;
    c->presieve_chunk = 1u << (segment_bits - 12); /* 4096 odd numbers per segment */
    return WW_OK;
}

int ww_get_launch_shape(ww_ctx *c, uint32_t *waves_per_group, uint32_t *segment_bits)
{
    WW_LOCK(c);
    if (!c) return fail(WW_E_ARG, "null context");
    if (waves_per_group) *waves_per_group = (uint32_t)c->waves;
    if (segment_bits) *segment_bits = 12 + (uint32_t)__builtin_ctz(c->presieve_chunk);
    return WW_OK;
}

int ww_clear_results(ww_ctx *c)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    /* checksum, totalprimecount, then primecount_max + resultcount (the list length
       primecount is cleared by ww_clear_prp_count only) */
    HIP_TRY(hipMemsetAsync(&c->d_state->checksum, 0, 16, c->stream));
    HIP_TRY(hipMemsetAsync(&c->d_state->primecount_max, 0, 8, c->stream));
    return WW_OK;
}

int ww_clear_prp_count(ww_ctx *c)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    HIP_TRY(hipMemsetAsync(&c->d_state->primecount, 0, 4, c->stream));
    return WW_OK;
}

int ww_getsegprimes(ww_ctx *c, uint64_t low, uint64_t high)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (low < WW_MIN_START) return fail(WW_E_ARG, "low %llu below %llu", (unsigned long long)low, WW_MIN_START);
    if (high <= low) return fail(WW_E_ARG, "empty range");
    if (high - low > (1ull << 34)) return fail(WW_E_ARG, "range wider than 2^34");
    hipEvent_t a, b;
    if ((rc = event_pair(c, c->ev_sieve, &a, &b))) return rc;
    HIP_TRY(hipEventRecord(a, c->stream));
    const uint64_t *pre = nullptr, *prer = nullptr;
    if ((rc = chunk_buckets(c, low, high, &pre, &prer))) return rc;
    if (!prer && (rc = chunk_presieve(c, low, high, &pre))) return rc;
    HIP_TRY(ww::launch_sieve_prp(low, high, c->d_list, c->list_cap, c->d_state, c->d_deep,
                                 c->ndeep, pre, c->stream, prer));
    HIP_TRY(hipEventRecord(b, c->stream));
    c->launches++;
    if (c->ev_sieve.size() >= kMaxPendingEvents) return collect(c->ev_sieve, &c->ms_sieve);
    return WW_OK;
}

int ww_wieferich_wallsunsun(ww_ctx *c)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    hipEvent_t a, b;
    if ((rc = event_pair(c, c->ev_tests, &a, &b))) return rc;
    HIP_TRY(hipEventRecord(a, c->stream));
    HIP_TRY(ww::launch_ww_tests(c->d_list, c->list_cap, c->d_state, c->d_results, c->result_cap,
                                c->test_grid, c->stream));
    HIP_TRY(hipEventRecord(b, c->stream));
    if (c->ev_tests.size() >= kMaxPendingEvents) return collect(c->ev_tests, &c->ms_tests);
    return WW_OK;
}

int ww_search_range(ww_ctx *c, uint64_t low, uint64_t high, uint64_t chunk)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (low < WW_MIN_START) return fail(WW_E_ARG, "low below %llu", WW_MIN_START);
    if (high <= low) return fail(WW_E_ARG, "empty range");
    if (chunk == 0) chunk = c->default_chunk;
    if (chunk > (1ull << 34)) chunk = 1ull << 34;
    for (uint64_t s = low; s < high;) {
        uint64_t e = s + chunk;
        if (e > high || e < s) e = high;
        if (ww_list_bound(s, e) > c->list_cap)
            return fail(WW_E_CAPACITY, "chunk [%llu,%llu) may exceed the list capacity %llu",
                        (unsigned long long)s, (unsigned long long)e, (unsigned long long)c->list_cap);
        if ((rc = ww_clear_prp_count(c))) return rc;
        if ((rc = ww_getsegprimes(c, s, e))) return rc;
        if ((rc = ww_wieferich_wallsunsun(c))) return rc;
        s = e;
    }
    return WW_OK;
}

int ww_search_fused(ww_ctx *c, uint64_t low, uint64_t high)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (low < WW_MIN_START) return fail(WW_E_ARG, "low %llu below %llu", (unsigned long long)low, WW_MIN_START);
    if (high <= low) return fail(WW_E_ARG, "empty range");
    if (high - low > (1ull << 34)) return fail(WW_E_ARG, "range wider than 2^34");
    if (!c->d_fstate) {
        hipError_t e = hipMalloc(&c->d_fsum, sizeof(SlotSum));
        if (e == hipSuccess) e = hipMalloc(&c->d_fres, (size_t)c->result_cap * sizeof(ww_result));
        if (e == hipSuccess) e = hipMalloc(&c->d_fstate, sizeof(ww::BatchState)); /* last: marks all made */
        if (e != hipSuccess) {
            if (c->d_fsum) (void)hipFree(c->d_fsum);
            if (c->d_fres) (void)hipFree(c->d_fres);
            c->d_fsum = nullptr;
            c->d_fres = nullptr;
            c->d_fstate = nullptr;
            return fail(WW_E_HIP, "ww_search_fused: %s", hipGetErrorString(e));
        }
    }
    hipEvent_t a, b;
    if ((rc = event_pair(c, c->ev_tests, &a, &b))) return rc;
    HIP_TRY(hipEventRecord(a, c->stream));
    const uint64_t *pre = nullptr, *prer = nullptr;
    if ((rc = chunk_buckets(c, low, high, &pre, &prer))) return rc;
    if (!prer && (rc = chunk_presieve(c, low, high, &pre))) return rc;
    int nb = c->tuned_blocks > 0 ? c->tuned_blocks : ww::batch_blocks_per_cu(c->waves);
    if (nb <= 0) nb = 4;
    HIP_TRY(ww::launch_batch(c->d_span1, c->d_prefix1, 1, c->d_fstate, c->d_fsum, 1, c->d_fres,
                             c->result_cap, c->cus * nb, c->tuned_claim, c->d_deep, c->ndeep, pre,
                             c->stream, c->waves, prer));
    HIP_TRY(ww::launch_fold(c->d_fstate, c->d_fsum, c->d_fres, c->result_cap, c->d_state,
                            c->d_results, c->result_cap, c->stream));
    HIP_TRY(hipEventRecord(b, c->stream));
    c->launches++;
    if (c->ev_tests.size() >= kMaxPendingEvents) return collect(c->ev_tests, &c->ms_tests);
    return WW_OK;
}

int ww_search_ranges_noalloc(ww_ctx *c, const ww_range *d_ranges, uint32_t n,
                             ww_range_sum *d_sums, ww_result *d_records, uint32_t record_cap,
                             uint32_t *d_record_count)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (n == 0 || n > WW_NOALLOC_MAX_RANGES)
        return fail(WW_E_CAPACITY, "ww_search_ranges_noalloc: %u ranges (1..%u allowed)", n,
                    WW_NOALLOC_MAX_RANGES);
    if (!d_ranges || !d_sums || !d_record_count || (record_cap && !d_records))
        return fail(WW_E_ARG, "ww_search_ranges_noalloc: null buffer");
    int nb = c->tuned_blocks > 0 ? c->tuned_blocks : ww::batch_blocks_per_cu(c->waves);
    if (nb <= 0) nb = 4;
    HIP_TRY(ww::launch_noalloc(c->device, d_ranges, n, reinterpret_cast<SlotSum *>(d_sums),
                               d_records, record_cap, d_record_count, c->cus * nb, c->tuned_claim,
                               c->d_deep, c->ndeep, c->stream, c->waves));
    return WW_OK;
}

int ww_finish(ww_ctx *c)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return WW_OK;
}

int ww_read_counters(ww_ctx *c, ww_counters *out)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!out) return fail(WW_E_ARG, "null out");
    DevState s;
    HIP_TRY(hipMemcpyAsync(&s, c->d_state, sizeof s, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    out->checksum = s.checksum;
    out->totalprimecount = s.totalprimecount;
    out->primecount = s.primecount;
    out->primecount_max = s.primecount_max;
    out->resultcount = s.resultcount;
    out->reserved = 0;
    return WW_OK;
}

int ww_read_results(ww_ctx *c, ww_result *out, uint32_t cap, uint32_t *n)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    ww_counters k;
    if ((rc = ww_read_counters(c, &k))) return rc;
    uint32_t avail = k.resultcount < c->result_cap ? k.resultcount : c->result_cap;
    uint32_t m = avail < cap ? avail : cap;
    if (m) {
        if (!out) return fail(WW_E_ARG, "null out");
        HIP_TRY(hipMemcpyAsync(out, c->d_results, (size_t)m * sizeof(ww_result),
                               hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (n) *n = m;
    return WW_OK;
}

int ww_read_prps(ww_ctx *c, uint64_t *out, uint64_t cap, uint64_t *n)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    ww_counters k;
    if ((rc = ww_read_counters(c, &k))) return rc;
    uint64_t len = k.primecount;
    uint64_t m = len < c->list_cap ? len : c->list_cap;
    if (m > cap) m = cap;
    if (m) {
        if (!out) return fail(WW_E_ARG, "null out");
        HIP_TRY(hipMemcpyAsync(out, c->d_list, m * sizeof(uint64_t), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (n) *n = len;
    return WW_OK;
}

int ww_kernel_times(ww_ctx *c, double *ms_sieve, double *ms_tests, uint64_t *launches)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if ((rc = collect(c->ev_sieve, &c->ms_sieve))) return rc;
    if ((rc = collect(c->ev_tests, &c->ms_tests))) return rc;
    if (ms_sieve) *ms_sieve = c->ms_sieve;
    if (ms_tests) *ms_tests = c->ms_tests;
    if (launches) *launches = c->launches;
    c->ms_sieve = c->ms_tests = 0;
    c->launches = 0;
    return WW_OK;
}

int ww_residues(ww_ctx *c, const uint64_t *p, uint64_t n, uint64_t *out)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (n == 0) return WW_OK;
    if (!p || !out) return fail(WW_E_ARG, "null buffer");
    for (uint64_t i = 0; i < n; ++i)
        if (p[i] < WW_MIN_START || (p[i] & 1) == 0 || p[i] % 5 == 0)
            return fail(WW_E_ARG, "p[%llu] = %llu is not odd, >= 127 and prime to 5",
                        (unsigned long long)i, (unsigned long long)p[i]);
    uint64_t *dp = nullptr, *dout = nullptr;
    HIP_TRY(hipMalloc(&dp, n * sizeof(uint64_t)));
    hipError_t e = hipMalloc(&dout, 4 * n * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpyAsync(dp, p, n * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = ww::launch_residues(dp, n, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, 4 * n * 8, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dp);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(WW_E_HIP, "ww_residues: %s", hipGetErrorString(e));
    return WW_OK;
}

int ww_prp_flags(ww_ctx *c, const uint64_t *n, uint64_t count, uint8_t *out)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (count == 0) return WW_OK;
    if (!n || !out) return fail(WW_E_ARG, "null buffer");
    for (uint64_t i = 0; i < count; ++i)
        if (n[i] < 3 || (n[i] & 1) == 0)
            return fail(WW_E_ARG, "n[%llu] = %llu is not odd and > 1", (unsigned long long)i,
                        (unsigned long long)n[i]);
    uint64_t *dn = nullptr;
    uint8_t *dout = nullptr;
    HIP_TRY(hipMalloc(&dn, count * sizeof(uint64_t)));
    hipError_t e = hipMalloc(&dout, count);
    if (e == hipSuccess) e = hipMemcpyAsync(dn, n, count * 8, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = ww::launch_prp_flags(dn, count, dout, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, count, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(dn);
    if (dout) (void)hipFree(dout);
    if (e != hipSuccess) return fail(WW_E_HIP, "ww_prp_flags: %s", hipGetErrorString(e));
    return WW_OK;
}

} // extern "C"

/* ---------------------------------------------------------------------------
 * Persistent form
 * ------------------------------------------------------------------------- */

struct ww_batch {
    int device = 0;
    hipStream_t stream = nullptr; /* the creating context's stream: every launch goes there */
    uint32_t n = 0;
    uint32_t res_cap = 0;
    int blocks_per_cu = 0;
    uint32_t claim = 8;
    BatchSpan *d_spans = nullptr;
    uint64_t *d_prefix = nullptr;
    SlotSum *d_sums = nullptr;
    BatchState *d_state = nullptr;
    ww_result *d_res = nullptr;
    uint64_t *d_pre = nullptr; /* wide pre-sieve words, 64 per segment (or none) */
    uint64_t nseg = 0;
    hipGraph_t graph = nullptr, graph_pre = nullptr;
    hipGraphExec_t exec = nullptr, exec_pre = nullptr;
    int graph_grid = 0;
    int graph_waves = 0;
    uint32_t graph_chunk = 0;
    uint32_t graph_claim = 0;
    uint32_t graph_ndeep = 0;
    uint64_t graph_wide = 0;
    const void *graph_wide_table = nullptr; /* the table the graph reads (rebuilds move it) */
    bool graph_buckets = false;
    hipEvent_t ev0 = nullptr, evm = nullptr, ev1 = nullptr;
    /* bucket sieve (ranges at least the context's bucket_min_width wide when the batch was
       made): the bucket-span table (host copy and device), the buckets, and the words of
       candidates that need the strong PRP test (64 per segment, like d_pre) */
    std::vector<ww::BucketSpan> bspans;
    ww::BucketSpan *d_bspans = nullptr;
    uint32_t nbands = 0;
    uint32_t *d_band_ent = nullptr, *d_band_cnt = nullptr;
    uint32_t band_cap = 0;
    uint32_t *d_bent = nullptr, *d_bcnt = nullptr;
    uint32_t bcap = 0;
    uint32_t *d_bover = nullptr; /* chunks whose buckets overflowed, summed over launches */
    uint64_t *d_prer = nullptr;
    uint64_t bucket_first = 0; /* table index the device's bucket-span table was made for */
};

namespace {

int batch_grid(ww_ctx *c, ww_batch *b)
{
    int nb = b->blocks_per_cu > 0 ? b->blocks_per_cu : ww::batch_blocks_per_cu(c->waves);
    if (nb <= 0) nb = 4;
    return c->cus * nb;
}

/* Whether this launch pre-sieves (the batch has the buffer and the context active
   pre-sieve primes). */
bool batch_presieves(ww_ctx *c, ww_batch *b)
{
    uint32_t first, count;
    wide_span(c, &first, &count);
    return count && b->d_pre;
}

/* Whether this launch uses the bucket sieve (known primes skip the strong PRP test): the
   batch has bucket spans, pre-sieves, and the context's sieve tiers allow it. */
bool batch_buckets(ww_ctx *c, ww_batch *b)
{
    return !b->bspans.empty() && b->d_prer && batch_presieves(c, b) && bucket_config_ok(c);
}

/* The bucket-span table for the context's pre-sieve bound (the sieving primes start above
   it): built on the host and copied with a wait, so never under capture. Builds the prime
   table on first use. */
int batch_prepare_buckets(ww_ctx *c, ww_batch *b)
{
    int rc = ensure_primes32(c);
    if (rc) return rc;
    const uint64_t first = bucket_first_index(c);
    if (first == b->bucket_first) return WW_OK;
    for (ww::BucketSpan &bs : b->bspans) {
        /* all table primes of the 2^21-integer blocks up to the top's; the kernel stops at
           the first prime above the top */
        const uint64_t blk = ((uint64_t)bs.top >> 21) + 1;
        bs.pidx0 = first;
        bs.pidx1 = c->primes32_prefix[std::min<uint64_t>(blk, c->primes32_prefix.size() - 1)];
        if (bs.pidx1 < bs.pidx0) bs.pidx1 = bs.pidx0;
    }
    HIP_TRY(hipMemcpyAsync(b->d_bspans, b->bspans.data(), b->bspans.size() * sizeof(ww::BucketSpan),
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    b->bucket_first = first;
    return WW_OK;
}

int batch_enqueue_presieve(ww_ctx *c, ww_batch *b, bool buckets)
{
    uint32_t first, count;
    wide_span(c, &first, &count);
    const uint64_t chunks = (b->nseg + c->presieve_chunk - 1) / c->presieve_chunk;
    const int pgrid = (int)std::min<uint64_t>(chunks, (uint64_t)c->cus);
    if (!buckets) {
        HIP_TRY(ww::launch_presieve(b->d_spans, b->d_prefix, b->n, c->d_wide + first, count,
                                    b->d_pre, pgrid > 0 ? pgrid : 1, c->stream, c->presieve_chunk));
        return WW_OK;
    }
    const ww::BucketStore st{b->d_band_ent, b->d_band_cnt, b->band_cap, b->d_bent, b->d_bcnt, b->bcap};
    HIP_TRY(ww::launch_bucket_sieve(b->d_spans, b->d_prefix, b->d_bspans, (uint32_t)b->bspans.size(),
                                    b->nbands, c->d_primes32, c->d_orders32, st, c->stream));
    ww::BucketArgs bk{b->d_bspans, b->d_bent, b->d_bcnt, b->bcap, b->d_bover};
    HIP_TRY(ww::launch_presieve(b->d_spans, b->d_prefix, b->n, c->d_wide + first, count, b->d_pre,
                                pgrid > 0 ? pgrid : 1, c->stream, c->presieve_chunk, b->d_prer, &bk));
    return WW_OK;
}

int batch_enqueue_search(ww_ctx *c, ww_batch *b, int grid, bool pre, bool buckets)
{
    HIP_TRY(ww::launch_batch(b->d_spans, b->d_prefix, b->n, b->d_state, b->d_sums, b->n, b->d_res,
                             b->res_cap, grid, b->claim, c->d_deep, c->ndeep,
                             pre ? b->d_pre : nullptr, c->stream, c->waves,
                             buckets ? b->d_prer : nullptr));
    return WW_OK;
}

/* Capture `enqueue` on the context's stream into an instantiated graph. */
template <class F> int capture(ww_ctx *c, hipGraph_t *graph, hipGraphExec_t *exec, F enqueue)
{
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return fail(WW_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    *graph = g;
    HIP_TRY(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
    return WW_OK;
}

void drop_graphs(ww_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->exec_pre) (void)hipGraphExecDestroy(b->exec_pre);
    if (b->graph_pre) (void)hipGraphDestroy(b->graph_pre);
    b->exec = b->exec_pre = nullptr;
    b->graph = b->graph_pre = nullptr;
}

} // namespace

extern "C" {

int ww_batch_create(ww_ctx *c, const ww_range *ranges, uint32_t n, uint32_t result_capacity,
                    ww_batch **out)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!out || !ranges || n == 0 || result_capacity == 0) return fail(WW_E_ARG, "bad batch arguments");
    *out = nullptr;
    std::vector<BatchSpan> spans(n);
    std::vector<uint64_t> prefix(n + 1, 0);
    for (uint32_t i = 0; i < n; ++i) {
        const ww_range &r = ranges[i];
        if (r.lo < WW_MIN_START || r.hi <= r.lo)
            return fail(WW_E_ARG, "range %u [%llu, %llu) is empty or starts below %llu", i,
                        (unsigned long long)r.lo, (unsigned long long)r.hi, WW_MIN_START);
        spans[i].lo = r.lo;
        spans[i].hi = r.hi;
        spans[i].seg0 = 0;
        spans[i].slot = i;
        spans[i].pad = 0;
        spans[i].pre_seg = prefix[i]; /* the batch's pre-sieve words follow its layout */
        prefix[i + 1] = prefix[i] + ww::segments_of(r.lo, r.hi);
    }
    ww_batch *b = new ww_batch;
    b->device = c->device;
    b->stream = c->stream;
    b->blocks_per_cu = c->tuned_blocks;
    b->claim = c->tuned_claim;
    b->n = n;
    b->res_cap = result_capacity;
    hipError_t e = hipMalloc(&b->d_spans, n * sizeof(BatchSpan));
    if (e == hipSuccess) e = hipMalloc(&b->d_prefix, (n + 1) * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc(&b->d_sums, n * sizeof(SlotSum));
    if (e == hipSuccess) e = hipMalloc(&b->d_state, sizeof(BatchState));
    if (e == hipSuccess) e = hipMalloc(&b->d_res, (size_t)result_capacity * sizeof(ww_result));
    b->nseg = prefix[n];
    if (e == hipSuccess && c->presieve_bound && b->nseg)
        e = hipMalloc(&b->d_pre, b->nseg * ww::kPresieveWordsPerSeg * sizeof(uint64_t));
    /* bucket spans: ranges at least bucket_min_width wide (and at most kBucketMaxChunks
       chunks). Buckets are sized for primes from 2^20 to sqrt(hi - 1): per 2^20-bit chunk,
       2^20 x kVisited (the odd multiples the walk visits) x sum(1/l) hits, sum(1/l) from Mertens'
       ln ln top - ln ln 2^20 (within 0.005 of the true sum). A chunk part's bucket (one of
       kBucketSplitParts) holds 1.15x its share of that mean plus 512; a band bucket (16
       chunks, one per fill workgroup) 1.25x its mean plus 256. An overflow only costs the
       chunks concerned their known primes. */
    if (e == hipSuccess && c->presieve_bound && b->nseg) {
        uint64_t chunks = 0, bands = 0;
        double mean = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const uint64_t nseg = prefix[i + 1] - prefix[i];
            if (ranges[i].hi - ranges[i].lo < c->bucket_min_width || nseg == 0) continue;
            const uint64_t bit0 = prefix[i] * 4096;
            const uint64_t cf = bit0 >> ww::kBucketChunkBits;
            const uint64_t nch = ((bit0 + nseg * 4096 - 1) >> ww::kBucketChunkBits) - cf + 1;
            if (nch > ww::kBucketMaxChunks) continue;
            ww::BucketSpan bs{};
            bs.span = i;
            bs.bit0 = bit0;
            bs.chunk_first = (uint32_t)cf;
            bs.nchunks = (uint32_t)nch;
            bs.nbands = (uint32_t)((nch + ww::kBucketBandChunks - 1) / ww::kBucketBandChunks);
            bs.top = (uint32_t)isqrt64(ranges[i].hi - 1);
            bs.cnt_base = chunks * ww::kBucketSplitParts;
            bs.band_cnt_base = bands * ww::kBucketGroups;
            chunks += nch;
            bands += bs.nbands;
            const double lt = std::log((double)bs.top), l20 = 20.0 * std::log(2.0);
            if (lt > l20) mean = std::max(mean, std::ldexp(1.0, 20) * kVisited * std::log(lt / l20));
            b->bspans.push_back(bs);
            spans[i].pad = (uint32_t)b->bspans.size();
        }
        if (!b->bspans.empty()) {
            const double sc = c->bucket_scale;
            b->bcap = (uint32_t)std::ceil(sc * (1.15 * mean / ww::kBucketSplitParts + 512));
            b->band_cap = (uint32_t)std::ceil(sc * (1.25 * mean * ww::kBucketBandChunks / ww::kBucketGroups + 256));
            b->nbands = (uint32_t)bands;
            for (ww::BucketSpan &bs : b->bspans) {
                bs.ent_base = bs.cnt_base * b->bcap;
                bs.band_ent_base = bs.band_cnt_base * b->band_cap;
            }
            const size_t nb = b->bspans.size();
            const size_t band_buckets = bands * ww::kBucketGroups;
            e = hipMalloc(&b->d_bspans, nb * sizeof(ww::BucketSpan));
            if (e == hipSuccess) e = hipMalloc(&b->d_band_cnt, band_buckets * sizeof(uint32_t));
            if (e == hipSuccess)
                e = hipMalloc(&b->d_band_ent, band_buckets * (size_t)b->band_cap * sizeof(uint32_t));
            const size_t parts = chunks * ww::kBucketSplitParts;
            if (e == hipSuccess) e = hipMalloc(&b->d_bcnt, parts * sizeof(uint32_t));
            if (e == hipSuccess) e = hipMalloc(&b->d_bent, parts * (size_t)b->bcap * sizeof(uint32_t));
            if (e == hipSuccess) e = hipMalloc(&b->d_bover, sizeof(uint32_t));
            if (e == hipSuccess) e = hipMemsetAsync(b->d_bover, 0, sizeof(uint32_t), c->stream);
            if (e == hipSuccess)
                e = hipMalloc(&b->d_prer, b->nseg * ww::kPresieveWordsPerSeg * sizeof(uint64_t));
        }
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_spans, spans.data(), n * sizeof(BatchSpan), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->d_prefix, prefix.data(), (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice,
                           c->stream);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->evm);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        ww_batch_destroy(b);
        return fail(WW_E_HIP, "ww_batch_create: %s", hipGetErrorString(e));
    }
    *out = b;
    return WW_OK;
}

int ww_batch_set_grid(ww_batch *b, int blocks_per_cu, uint32_t claim)
{
    if (!b || blocks_per_cu < 0 || claim > 4096) return fail(WW_E_ARG, "bad grid");
    b->blocks_per_cu = blocks_per_cu;
    b->claim = claim ? claim : 8;
    return WW_OK;
}

int ww_batch_launch(ww_ctx *c, ww_batch *b)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b) return fail(WW_E_ARG, "null batch");
    const bool pre = batch_presieves(c, b);
    const bool buckets = batch_buckets(c, b);
    if (buckets && (rc = batch_prepare_buckets(c, b))) return rc;
    HIP_TRY(hipEventRecord(b->ev0, c->stream));
    if (pre && (rc = batch_enqueue_presieve(c, b, buckets))) return rc;
    HIP_TRY(hipEventRecord(b->evm, c->stream));
    if ((rc = batch_enqueue_search(c, b, batch_grid(c, b), pre, buckets))) return rc;
    HIP_TRY(hipEventRecord(b->ev1, c->stream));
    return WW_OK;
}

int ww_batch_launch_graph(ww_ctx *c, ww_batch *b)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b) return fail(WW_E_ARG, "null batch");
    int grid = batch_grid(c, b);
    uint32_t wfirst, wcount;
    wide_span(c, &wfirst, &wcount);
    const uint64_t wide = ((uint64_t)wfirst << 32) | wcount;
    const bool pre = batch_presieves(c, b);
    const bool buckets = batch_buckets(c, b);
    if (buckets && (rc = batch_prepare_buckets(c, b))) return rc;
    if (b->exec && (b->graph_grid != grid || b->graph_claim != b->claim ||
                    b->graph_ndeep != c->ndeep || b->graph_wide != wide ||
                    b->graph_wide_table != c->d_wide || b->graph_waves != c->waves ||
                    b->graph_chunk != c->presieve_chunk || b->graph_buckets != buckets))
        drop_graphs(b);
    if (!b->exec) {
        if (pre && (rc = capture(c, &b->graph_pre, &b->exec_pre,
                                 [&] { return batch_enqueue_presieve(c, b, buckets); }))) {
            drop_graphs(b);
            return rc;
        }
        if ((rc = capture(c, &b->graph, &b->exec,
                          [&] { return batch_enqueue_search(c, b, grid, pre, buckets); }))) {
            drop_graphs(b);
            return rc;
        }
        b->graph_buckets = buckets;
        b->graph_grid = grid;
        b->graph_waves = c->waves;
        b->graph_chunk = c->presieve_chunk;
        b->graph_claim = b->claim;
        b->graph_ndeep = c->ndeep;
        b->graph_wide = wide;
        b->graph_wide_table = c->d_wide;
    }
    /* two graphs, so that the events between them time each kernel on its own */
    HIP_TRY(hipEventRecord(b->ev0, c->stream));
    if (b->exec_pre) HIP_TRY(hipGraphLaunch(b->exec_pre, c->stream));
    HIP_TRY(hipEventRecord(b->evm, c->stream));
    HIP_TRY(hipGraphLaunch(b->exec, c->stream));
    HIP_TRY(hipEventRecord(b->ev1, c->stream));
    return WW_OK;
}

int ww_batch_read(ww_ctx *c, ww_batch *b, ww_range_sum *sums, ww_result *recs, uint32_t cap,
                  uint32_t *nrec, uint32_t *total)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b) return fail(WW_E_ARG, "null batch");
    BatchState st;
    HIP_TRY(hipMemcpyAsync(&st, b->d_state, sizeof st, hipMemcpyDeviceToHost, c->stream));
    if (sums) {
        static_assert(sizeof(ww_range_sum) == sizeof(SlotSum), "layout");
        HIP_TRY(hipMemcpyAsync(sums, b->d_sums, b->n * sizeof(SlotSum), hipMemcpyDeviceToHost, c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    uint32_t avail = st.resultcount < b->res_cap ? st.resultcount : b->res_cap;
    uint32_t m = recs ? (avail < cap ? avail : cap) : 0;
    if (m) {
        HIP_TRY(hipMemcpyAsync(recs, b->d_res, (size_t)m * sizeof(ww_result), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (nrec) *nrec = m;
    if (total) *total = st.resultcount;
    return WW_OK;
}

int ww_batch_time(ww_ctx *c, ww_batch *b, double *ms)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b || !ms) return fail(WW_E_ARG, "null argument");
    float t = 0;
    HIP_TRY(hipEventSynchronize(b->ev1));
    HIP_TRY(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t;
    return WW_OK;
}

int ww_batch_kernel_times(ww_ctx *c, ww_batch *b, double *presieve_ms, double *search_ms)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b) return fail(WW_E_ARG, "null batch");
    float t0 = 0, t1 = 0;
    HIP_TRY(hipEventSynchronize(b->ev1));
    HIP_TRY(hipEventElapsedTime(&t0, b->ev0, b->evm));
    HIP_TRY(hipEventElapsedTime(&t1, b->evm, b->ev1));
    if (presieve_ms) *presieve_ms = t0;
    if (search_ms) *search_ms = t1;
    return WW_OK;
}

int ww_batch_buckets(ww_ctx *c, ww_batch *b, ww_bucket_info *out)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!b || !out) return fail(WW_E_ARG, "null argument");
    memset(out, 0, sizeof *out);
    out->spans = (uint32_t)b->bspans.size();
    out->active = batch_buckets(c, b) ? 1u : 0u;
    out->bucket_cap = b->bcap;
    uint64_t chunks = 0;
    for (const ww::BucketSpan &bs : b->bspans) chunks += bs.nchunks;
    out->bytes = (chunks * ww::kBucketSplitParts * ((uint64_t)b->bcap + 1) +
                  (uint64_t)b->nbands * ww::kBucketGroups * ((uint64_t)b->band_cap + 1)) *
                 sizeof(uint32_t);
    if (b->d_bover) {
        HIP_TRY(hipMemcpyAsync(&out->overflows, b->d_bover, sizeof(uint32_t), hipMemcpyDeviceToHost,
                               c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return WW_OK;
}

int ww_batch_destroy(ww_batch *b)
{
    if (!b) return WW_OK;
    (void)hipSetDevice(b->device);
    /* launches of this batch are queued on its context's stream only: waiting for that
       stream leaves other contexts and servers on the device running */
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->evm) (void)hipEventDestroy(b->evm);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->d_pre) (void)hipFree(b->d_pre);
    if (b->d_prer) (void)hipFree(b->d_prer);
    if (b->d_bspans) (void)hipFree(b->d_bspans);
    if (b->d_bent) (void)hipFree(b->d_bent);
    if (b->d_bcnt) (void)hipFree(b->d_bcnt);
    if (b->d_band_ent) (void)hipFree(b->d_band_ent);
    if (b->d_band_cnt) (void)hipFree(b->d_band_cnt);
    if (b->d_bover) (void)hipFree(b->d_bover);
    if (b->d_res) (void)hipFree(b->d_res);
    if (b->d_state) (void)hipFree(b->d_state);
    if (b->d_sums) (void)hipFree(b->d_sums);
    if (b->d_prefix) (void)hipFree(b->d_prefix);
    if (b->d_spans) (void)hipFree(b->d_spans);
    delete b;
    return WW_OK;
}

int ww_autotune(ww_ctx *c, uint64_t lo, uint64_t hi, const ww_tune_space *space,
                ww_tuning *best, ww_tuning *trials, uint32_t cap, uint32_t *ntrials)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    static const uint32_t kBounds[] = {0, 1024, 2048, 4096, 8192, 16384, 65535};
    static const int32_t kBlocks[] = {0, 4, 5};
    static const uint32_t kClaims[] = {1, 4, 16, 64};
    static const uint32_t kPresieve[] = {0, 1u << 18, 1u << 19, 1u << 20, 1u << 21};
    static const uint32_t kWaves[] = {2, 4, 8};
    static const uint32_t kSegBits[] = {15, 16, 17, 18, 19, 20};
    ww_tune_space sp{};
    if (space) sp = *space;
#define WW_DEFAULT_LIST(f, n, d)                                                         \
    if (!sp.f || !sp.n) {                                                                \
        sp.f = d;                                                                        \
        sp.n = sizeof d / sizeof *d;                                                     \
    }
    WW_DEFAULT_LIST(deep_bounds, n_deep_bounds, kBounds)
    WW_DEFAULT_LIST(blocks_per_cu, n_blocks_per_cu, kBlocks)
    WW_DEFAULT_LIST(claims, n_claims, kClaims)
    WW_DEFAULT_LIST(presieve_bounds, n_presieve_bounds, kPresieve)
    WW_DEFAULT_LIST(waves_per_group, n_waves_per_group, kWaves)
    WW_DEFAULT_LIST(segment_bits, n_segment_bits, kSegBits)
#undef WW_DEFAULT_LIST
    for (uint32_t i = 0; i < sp.n_presieve_bounds; ++i)
        if (sp.presieve_bounds[i] > WW_MAX_PRESIEVE_BOUND)
            return fail(WW_E_ARG, "pre-sieve bound %u above %u", sp.presieve_bounds[i], WW_MAX_PRESIEVE_BOUND);
    for (uint32_t i = 0; i < sp.n_waves_per_group; ++i)
        if (!ww::batch_waves_ok((int)sp.waves_per_group[i]))
            return fail(WW_E_ARG, "waves per workgroup %u: 2, 4 or 8", sp.waves_per_group[i]);
    for (uint32_t i = 0; i < sp.n_segment_bits; ++i)
        if (sp.segment_bits[i] < 15 || sp.segment_bits[i] > 20)
            return fail(WW_E_ARG, "segment bits %u outside 15..20", sp.segment_bits[i]);
    const uint32_t saved_bound = c->deep_bound, saved_pre = c->presieve_bound;
    const int saved_waves = c->waves;
    const uint32_t saved_chunk = c->presieve_chunk;
    auto restore = [&] {
        ww_set_deep_bound(c, saved_bound);
        ww_set_presieve_bound(c, saved_pre);
        c->waves = saved_waves;
        c->presieve_chunk = saved_chunk;
    };
    /* the sample batch holds pre-sieve words only if the pre-sieve is on when it is made */
    if (!saved_pre && (rc = ww_set_presieve_bound(c, WW_DEFAULT_PRESIEVE_BOUND))) return rc;
    ww_range r = {lo, hi};
    ww_batch *b = nullptr;
    if ((rc = ww_batch_create(c, &r, 1, 1u << 16, &b))) {
        restore();
        return rc;
    }
    if ((rc = ww_set_presieve_bound(c, saved_pre))) {
        ww_batch_destroy(b);
        restore();
        return rc;
    }
    const uint32_t seg_bits0 = 12 + (uint32_t)__builtin_ctz(c->presieve_chunk);
    ww_tuning top;
    memset(&top, 0, sizeof top);
    top.ms = 1e300;
    uint32_t n = 0;
    /* one timed configuration of the context's current bounds and shape (the first launch
       warms up) */
    auto trial = [&](int32_t bpc, uint32_t claim) -> int {
        int e;
        if ((e = ww_batch_set_grid(b, bpc, claim))) return e;
        double ms = 0, t = 0;
        ww_range_sum sum;
        for (int rep = 0; rep < 2; ++rep) {
            if ((e = ww_batch_launch(c, b))) return e;
            if ((e = ww_batch_time(c, b, &t))) return e;
            ms = t;
        }
        if ((e = ww_batch_read(c, b, &sum, nullptr, 0, nullptr, nullptr))) return e;
        ww_tuning tr;
        memset(&tr, 0, sizeof tr);
        tr.deep_bound = c->deep_bound;
        tr.blocks_per_cu = bpc;
        tr.claim = claim;
        tr.presieve_bound = c->presieve_bound;
        tr.ms = ms;
        tr.prps_per_s = ms > 0 ? sum.prp_count / (ms * 1e-3) : 0;
        tr.waves_per_group = (uint32_t)c->waves;
        tr.segment_bits = 12 + (uint32_t)__builtin_ctz(c->presieve_chunk);
        if (trials && n < cap) trials[n] = tr;
        ++n;
        if (ms < top.ms) top = tr;
        return WW_OK;
    };
    auto apply = [&](const ww_tuning &t) -> int {
        int e;
        if ((e = ww_set_deep_bound(c, t.deep_bound))) return e;
        if ((e = ww_set_presieve_bound(c, t.presieve_bound))) return e;
        return ww_set_launch_shape(c, t.waves_per_group, t.segment_bits);
    };
    /* stage 1: deep bound x geometry at the context's shape and pre-sieve bound */
    for (uint32_t i = 0; i < sp.n_deep_bounds && rc == WW_OK; ++i) {
        if ((rc = ww_set_deep_bound(c, sp.deep_bounds[i]))) break;
        for (uint32_t j = 0; j < sp.n_blocks_per_cu && rc == WW_OK; ++j)
            for (uint32_t k = 0; k < sp.n_claims && rc == WW_OK; ++k)
                rc = trial(sp.blocks_per_cu[j], sp.claims[k]);
    }
    /* stage 2: work-group shape; workgroups per CU scaled to it (the listed counts are for
       four-wave workgroups: b 4-wave groups hold as many waves as 4b/w groups of w waves) */
    if (rc == WW_OK) rc = apply(top);
    const ww_tuning s1 = top;
    for (uint32_t i = 0; i < sp.n_waves_per_group && rc == WW_OK; ++i) {
        const uint32_t w = sp.waves_per_group[i];
        if ((rc = ww_set_launch_shape(c, w, seg_bits0))) break;
        std::vector<int32_t> seen;
        for (uint32_t j = 0; j < sp.n_blocks_per_cu && rc == WW_OK; ++j) {
            const int32_t b4 = sp.blocks_per_cu[j];
            if (b4 && (4 * b4) % (int32_t)w) continue; /* no whole number of groups */
            const int32_t bw = b4 ? 4 * b4 / (int32_t)w : 0;
            if (std::find(seen.begin(), seen.end(), bw) != seen.end()) continue;
            seen.push_back(bw);
            if ((int)w == s1.waves_per_group && bw == s1.blocks_per_cu) continue; /* measured */
            rc = trial(bw, s1.claim);
        }
    }
    /* stage 3: pre-sieve bound; stage 4: the pre-sieve's segment size */
    if (rc == WW_OK) rc = apply(top);
    const ww_tuning s2 = top;
    for (uint32_t i = 0; i < sp.n_presieve_bounds && rc == WW_OK; ++i) {
        if (sp.presieve_bounds[i] == s2.presieve_bound) continue;
        if ((rc = ww_set_presieve_bound(c, sp.presieve_bounds[i]))) break;
        rc = trial(s2.blocks_per_cu, s2.claim);
    }
    if (rc == WW_OK) rc = apply(top);
    const ww_tuning s3 = top;
    for (uint32_t i = 0; i < sp.n_segment_bits && rc == WW_OK && s3.presieve_bound; ++i) {
        if (sp.segment_bits[i] == s3.segment_bits) continue;
        if ((rc = ww_set_launch_shape(c, s3.waves_per_group, sp.segment_bits[i]))) break;
        rc = trial(s3.blocks_per_cu, s3.claim);
    }
    ww_batch_destroy(b);
    if (rc == WW_OK) rc = apply(top);
    if (rc) {
        restore();
        return rc;
    }
    c->tuned_blocks = top.blocks_per_cu;
    c->tuned_claim = top.claim;
    c->last_tuning = top;
    if (best) *best = top;
    if (ntrials) *ntrials = n;
    return WW_OK;
}

int ww_tuning_save(ww_ctx *c, const char *path)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!path) return fail(WW_E_ARG, "null path");
    ww_device_info_t info;
    if ((rc = ww_device_info(c->device, &info))) return rc;
    FILE *f = fopen(path, "w");
    if (!f) return fail(WW_E_ARG, "cannot write %s", path);
    uint32_t waves = 0, bits = 0;
    ww_get_launch_shape(c, &waves, &bits);
    const int n = fprintf(f,
                          "{\n  \"device\": \"%s\",\n  \"arch\": \"%s\",\n  \"compute_units\": %d,\n"
                          "  \"deep_bound\": %u,\n  \"blocks_per_cu\": %d,\n  \"claim\": %u,\n"
                          "  \"presieve_bound\": %u,\n  \"waves_per_group\": %u,\n"
                          "  \"segment_bits\": %u,\n  \"ms\": %.6f,\n  \"prps_per_s\": %.1f\n}\n",
                          info.name, info.arch, info.compute_units, c->deep_bound, c->tuned_blocks,
                          c->tuned_claim, c->presieve_bound, waves, bits, c->last_tuning.ms,
                          c->last_tuning.prps_per_s);
    if (fclose(f) != 0 || n < 0) return fail(WW_E_ARG, "cannot write %s", path);
    return WW_OK;
}

} // extern "C"

namespace {

/* The value text after "key": in a flat JSON object (no nesting), or null. */
const char *json_value(const std::string &text, const char *key)
{
    const std::string k = std::string("\"") + key + "\"";
    size_t at = text.find(k);
    if (at == std::string::npos) return nullptr;
    at = text.find(':', at + k.size());
    if (at == std::string::npos) return nullptr;
    const char *v = text.c_str() + at + 1;
    while (*v == ' ' || *v == '\t' || *v == '\n' || *v == '\r') ++v;
    return v;
}

bool json_uint(const std::string &text, const char *key, long long lo, long long hi, long long *out)
{
    const char *v = json_value(text, key);
    if (!v) return false;
    char *end = nullptr;
    const long long x = strtoll(v, &end, 10);
    if (end == v || x < lo || x > hi) return false;
    *out = x;
    return true;
}

bool json_string(const std::string &text, const char *key, std::string *out)
{
    const char *v = json_value(text, key);
    if (!v || *v != '"') return false;
    const char *e = strchr(v + 1, '"');
    if (!e) return false;
    out->assign(v + 1, e);
    return true;
}

} // namespace

extern "C" {

int ww_tuning_load(ww_ctx *c, const char *path, ww_tuning *applied)
{
    WW_LOCK(c);
    int rc = use(c);
    if (rc) return rc;
    if (!path) return fail(WW_E_ARG, "null path");
    FILE *f = fopen(path, "r");
    if (!f) return fail(WW_E_ARG, "cannot read %s", path);
    std::string text;
    char buf[1024];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0 && text.size() < 65536) text.append(buf, k);
    fclose(f);
    ww_device_info_t info;
    if ((rc = ww_device_info(c->device, &info))) return rc;
    std::string name, arch;
    long long cus = 0, deep = 0, bpc = 0, claim = 0, pre = 0, waves = 0, bits = 0;
    if (!json_string(text, "device", &name) || !json_string(text, "arch", &arch) ||
        !json_uint(text, "compute_units", 1, 1 << 20, &cus) ||
        !json_uint(text, "deep_bound", 0, ww::kMaxDeepBound, &deep) ||
        !json_uint(text, "blocks_per_cu", 0, 64, &bpc) || !json_uint(text, "claim", 1, 4096, &claim) ||
        !json_uint(text, "presieve_bound", 0, WW_MAX_PRESIEVE_BOUND, &pre) ||
        !json_uint(text, "waves_per_group", 2, 8, &waves) || !json_uint(text, "segment_bits", 15, 20, &bits))
        return fail(WW_E_ARG, "%s is not a tuning file of this library", path);
    if (name != info.name || arch != info.arch || cus != info.compute_units)
        return fail(WW_E_ARG, "%s was tuned on %s (%s, %lld CUs), not on this %s (%s, %d CUs)", path,
                    name.c_str(), arch.c_str(), cus, info.name, info.arch, info.compute_units);
    if (deep && deep < 127) return fail(WW_E_ARG, "%s: deep bound %lld", path, deep);
    if ((rc = ww_set_deep_bound(c, (uint32_t)deep))) return rc;
    if ((rc = ww_set_presieve_bound(c, (uint32_t)pre))) return rc;
    if ((rc = ww_set_launch_shape(c, (uint32_t)waves, (uint32_t)bits))) return rc;
    c->tuned_blocks = (int)bpc;
    c->tuned_claim = (uint32_t)claim;
    ww_tuning t;
    memset(&t, 0, sizeof t);
    t.deep_bound = (uint32_t)deep;
    t.blocks_per_cu = (int32_t)bpc;
    t.claim = (uint32_t)claim;
    t.presieve_bound = (uint32_t)pre;
    t.waves_per_group = (uint32_t)waves;
    t.segment_bits = (uint32_t)bits;
    const char *v = json_value(text, "ms");
    if (v) t.ms = strtod(v, nullptr);
    v = json_value(text, "prps_per_s");
    if (v) t.prps_per_s = strtod(v, nullptr);
    c->last_tuning = t;
    if (applied) *applied = t;
    return WW_OK;
}

} // extern "C"
