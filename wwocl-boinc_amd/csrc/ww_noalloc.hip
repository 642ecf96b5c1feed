/*
 * ww_noalloc.hip -- the persistent search of many ranges with no allocation in the call.
 *
 * The batched form (ww_batch_*) allocates its span table, prefix, state, sums, records
 * and pre-sieve words when a batch is created. This variant takes device buffers for the
 * ranges (input) and the sums and records (outputs) from the caller and keeps what the
 * launch needs besides (the span table, the segment prefix and the persistent kernel's
 * state) in one file-scope device array, `scratch`, of fixed size. It runs without the
 * wide pre-sieve, whose words (512 bytes per 8192 integers) would not fit a fixed size;
 * results are the same, the launch is slower.
 *
 * `scratch` names a host-side stand-in in host code: its device address comes from
 * hipGetSymbolAddress, fetched once per device at set-up (ww_open) and used from then on.
 * Every call on a device shares that device's array, and calls that overlap on one device
 * (from contexts on different streams) race; one stream's calls are ordered.
 */
#include <mutex>

#include "ww_batch.h"

namespace ww {

namespace {

constexpr uint32_t kMaxRanges = WW_NOALLOC_MAX_RANGES;
constexpr size_t kStateOff = 0;
constexpr size_t kPrefixOff = 256;
constexpr size_t kSpansOff = (kPrefixOff + (kMaxRanges + 1) * sizeof(uint64_t) + 255) / 256 * 256;
constexpr size_t kScratchBytes = kSpansOff + kMaxRanges * sizeof(BatchSpan);
static_assert(sizeof(BatchState) <= kPrefixOff, "state fits its slot");
static_assert(sizeof(SlotSum) == sizeof(ww_range_sum), "sums are the caller's ww_range_sum");

/* The launch's scratch on each device: BatchState | prefix[kMaxRanges + 1] | spans[kMaxRanges]. */
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices]; /* device address of `scratch`, per device */

constexpr int kSpanThreads = 1024;
constexpr uint32_t kPerThread = kMaxRanges / kSpanThreads; /* 4 */

/* Span table and segment prefix of ranges[0..n) (n <= kMaxRanges). A range that is not
   WW_MIN_START <= lo < hi gets no segments (searched as empty, its sums stay zero). One
   workgroup: thread t owns ranges [4t, 4t + 4); their segment counts are scanned across
   the workgroup in LDS. */
__global__ __launch_bounds__(kSpanThreads) void k_noalloc_spans(const ww_range *ranges, uint32_t n,
                                                                BatchSpan *spans, uint64_t *prefix)
{
    __shared__ uint64_t part[kSpanThreads];
    const uint32_t t = threadIdx.x;
    uint64_t segs[kPerThread];
    uint64_t mine = 0;
    for (uint32_t k = 0; k < kPerThread; ++k) {
        const uint32_t i = t * kPerThread + k;
        segs[k] = 0;
        if (i < n) {
            const ww_range r = ranges[i];
            const bool ok = r.lo >= WW_MIN_START && r.hi > r.lo;
            segs[k] = ok ? segments_of(r.lo, r.hi) : 0;
            BatchSpan sp;
            sp.lo = ok ? r.lo : WW_MIN_START;
            sp.hi = ok ? r.hi : WW_MIN_START;
            sp.seg0 = 0;
            sp.slot = i;
            sp.pad = 0;
            sp.pre_seg = 0; /* no pre-sieve in this launch */
            spans[i] = sp;
        }
        mine += segs[k];
    }
    part[t] = mine;
    __syncthreads();
    for (uint32_t d = 1; d < kSpanThreads; d <<= 1) { /* inclusive scan */
        const uint64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    uint64_t run = part[t] - mine; /* exclusive prefix of this thread's first range */
    for (uint32_t k = 0; k < kPerThread; ++k) {
        const uint32_t i = t * kPerThread + k;
        if (i < n) prefix[i] = run;
        run += segs[k];
    }
    if (t == kSpanThreads - 1) prefix[n] = part[t];
}

__global__ void k_noalloc_count(const BatchState *st, uint32_t *count)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = st->resultcount;
}

} // namespace

hipError_t noalloc_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
    return hipSuccess;
}

hipError_t launch_noalloc(int device, const ww_range *ranges, uint32_t n, SlotSum *sums,
                          ww_result *res, uint32_t res_cap, uint32_t *res_count, int grid,
                          uint32_t claim, const DeepPrime *deep, uint32_t ndeep, hipStream_t s,
                          int waves)
{
    if (n == 0 || n > kMaxRanges) return hipErrorInvalidValue;
    void *base = nullptr;
    {
        std::lock_guard<std::mutex> g(g_mu);
        if (device >= 0 && device < kMaxDevices) base = g_ptr[device];
    }
    if (!base) return hipErrorNotInitialized;
    unsigned char *b = static_cast<unsigned char *>(base);
    BatchState *st = reinterpret_cast<BatchState *>(b + kStateOff);
    uint64_t *prefix = reinterpret_cast<uint64_t *>(b + kPrefixOff);
    BatchSpan *spans = reinterpret_cast<BatchSpan *>(b + kSpansOff);
    hipLaunchKernelGGL(k_noalloc_spans, dim3(1), dim3(kSpanThreads), 0, s, ranges, n, spans, prefix);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    /* clears the state and the caller's n sums, then the persistent kernel (no pre-sieve) */
    e = launch_batch(spans, prefix, n, st, sums, n, res, res_cap, grid, claim, deep, ndeep, nullptr, s,
                     waves);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_noalloc_count, dim3(1), dim3(64), 0, s, st, res_count);
    return hipGetLastError();
}

} // namespace ww
