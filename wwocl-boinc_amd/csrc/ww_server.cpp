/*
 * ww_server.cpp -- the serving form: continuous batching of range requests.
 *
 * An engine thread owns a HIP stream and a fixed pool of device slots (per-slot checksum
 * and count accumulate on the device across launches). Each admitted request keeps its
 * sieve state resident in HBM until it completes: the wide pre-sieve's words for all its
 * segments (512 bytes per 8192 integers), written once at admission into a block of the
 * server's arena, which is the HBM budget. Requests wait in an admission queue (FIFO)
 * until a slot and a free arena block of their size are available, so the budget bounds
 * how many requests, and how wide ones, are resident at once. Every "tick" the engine
 * takes the next segments of the active requests in admission order, up to the tick
 * size, and runs one persistent k_batch launch over them, each span reading its words
 * from its request's block; up to kRing ticks are in flight (span tables, record buffers
 * and sum snapshots are triple-buffered), so the host builds tick t+1 while tick t runs
 * and requests arriving meanwhile join the next tick. A request is complete when the
 * tick holding its last segment has finished (ticks retire in order, so no launch still
 * reads its block, which is then freed); its checksum and count come from that tick's
 * snapshot of the slot sums, its records from the ticks that produced them. Candidates
 * and PRPs stay in LDS: requests need no PRP buffers in HBM.
 *
 * No counterpart in the reference (one range per process, main.cpp:1110-1803); per
 * request the results equal the plain form's on that range.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/ww_api.h"
#include "ww_batch.h"

using ww::BatchSpan;
using ww::BatchState;
using ww::SlotSum;

namespace {

constexpr int kRing = 3;                    /* ticks in flight */
constexpr uint32_t kTickRecords = 1u << 16; /* result records per tick (capacity) */
constexpr uint32_t kTickCopy = 1024;        /* records copied back with every tick */
constexpr uint64_t kBlockBytes = ww::kPresieveWordsPerSeg * sizeof(uint64_t); /* 512: one segment's words */

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Req {
    uint64_t id, lo, hi, nseg = 0, next = 0;
    uint64_t checksum = 0, prp_count = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint32_t slot = 0;
    uint64_t blocks = 0;  /* arena blocks reserved (its sieve state + bytes_per_request) */
    uint64_t pre_off = 0; /* first block of its resident pre-sieve words */
    std::vector<ww_result> recs;
};

struct Tick {
    BatchSpan *d_spans = nullptr, *h_spans = nullptr;
    uint64_t *d_prefix = nullptr, *h_prefix = nullptr;
    BatchState *d_state = nullptr, *h_state = nullptr;
    ww_result *d_res = nullptr, *h_res = nullptr;
    SlotSum *h_sums = nullptr;
    hipEvent_t done = nullptr;
    bool busy = false;
    uint64_t segments = 0;
    std::vector<uint32_t> finishing;        /* slots whose last segment is in this tick */
    std::unordered_map<uint32_t, uint64_t> owner; /* slot -> request of this tick */
};

} // namespace

struct ww_server {
    ww_ctx *ctx = nullptr;
    int device = 0;
    ww_server_config cfg{};
    int grid = 0;
    hipStream_t stream = nullptr;
    SlotSum *d_sums = nullptr;
    /* the server's own copies of the context's active sieve tables, so that later changes
       to the context's bounds (which may free its tables) never reach a running engine */
    ww::DeepPrime *d_deep = nullptr;
    uint32_t ndeep = 0;
    ww::WidePrime *d_wide = nullptr;
    uint32_t nwide = 0;
    int cus = 0;
    int waves = ww::kDefaultBatchWaves;                     /* the context's shape at creation */
    uint32_t presieve_chunk = ww::kPresieveMaxChunkSegs;
    /* the HBM budget as an arena of 512-byte blocks: resident pre-sieve words of the
       admitted requests (allocated only when the pre-sieve is on; without it the blocks are
       only counted) */
    uint64_t *d_arena = nullptr;
    uint64_t arena_blocks = 0;
    std::map<uint64_t, uint64_t> free_blocks; /* first block -> length, coalesced */
    uint64_t extra_blocks = 0;                /* bytes_per_request in blocks */
    BatchSpan *d_adm_span = nullptr;          /* one-span table of an admission's pre-sieve */
    uint64_t *d_adm_prefix = nullptr;
    Tick ticks[kRing];
    int ring_next = 0;

    std::mutex mu;
    std::condition_variable cv;
    std::thread engine;
    bool stop = false;
    int error = 0;
    char errmsg[256] = "";
    uint64_t next_id = 1;
    std::deque<uint64_t> pending;            /* admission queue */
    std::unordered_map<uint64_t, Req> reqs;  /* submitted, not yet collected */
    std::vector<uint32_t> free_slots;
    std::vector<uint64_t> active;            /* admission order */
    std::deque<uint64_t> done;               /* completed, not yet polled */
    std::unordered_map<uint64_t, std::vector<ww_result>> taken; /* polled, records kept */
    ww_server_stats stats{};
};

namespace {

/* Blocks a request of nseg segments reserves. */
uint64_t blocks_for(const ww_server *s, uint64_t nseg)
{
    return (s->nwide ? nseg : 0) + s->extra_blocks;
}

/* First fit in the arena; false when no free run is long enough. Caller holds mu. */
bool arena_take(ww_server *s, uint64_t n, uint64_t *off)
{
    if (n == 0) {
        *off = 0;
        return true;
    }
    for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it) {
        if (it->second < n) continue;
        *off = it->first;
        const uint64_t rest = it->second - n, at = it->first + n;
        s->free_blocks.erase(it);
        if (rest) s->free_blocks.emplace(at, rest);
        return true;
    }
    return false;
}

/* Return a run to the arena, merged with its free neighbours. Caller holds mu. */
void arena_give(ww_server *s, uint64_t off, uint64_t n)
{
    if (n == 0) return;
    auto next = s->free_blocks.lower_bound(off);
    if (next != s->free_blocks.end() && off + n == next->first) {
        n += next->second;
        next = s->free_blocks.erase(next);
    }
    if (next != s->free_blocks.begin()) {
        auto prev = std::prev(next);
        if (prev->first + prev->second == off) {
            prev->second += n;
            return;
        }
    }
    s->free_blocks.emplace(off, n);
}

bool hip_ok(ww_server *s, hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    std::lock_guard<std::mutex> g(s->mu);
    if (!s->error) {
        s->error = WW_E_HIP;
        snprintf(s->errmsg, sizeof s->errmsg, "%s: %s", what, hipGetErrorString(e));
    }
    s->cv.notify_all();
    return false;
}

/* Finish a completed tick: records to their requests, finished requests to `done`. */
bool retire(ww_server *s, Tick &t)
{
    if (!hip_ok(s, hipEventSynchronize(t.done), "tick event")) return false;
    uint32_t nrec = t.h_state->resultcount;
    if (nrec > kTickRecords) {
        std::lock_guard<std::mutex> g(s->mu);
        s->error = WW_E_CAPACITY;
        snprintf(s->errmsg, sizeof s->errmsg, "%u records overflowed a tick", nrec);
        s->cv.notify_all();
        return false;
    }
    if (nrec > kTickCopy &&
        !hip_ok(s, hipMemcpy(t.h_res + kTickCopy, t.d_res + kTickCopy,
                             (nrec - kTickCopy) * sizeof(ww_result), hipMemcpyDeviceToHost),
                "records"))
        return false;
    double t_now = now_s();
    std::lock_guard<std::mutex> g(s->mu);
    for (uint32_t i = 0; i < nrec; ++i) {
        const ww_result &r = t.h_res[i];
        auto it = t.owner.find((uint32_t)r.pad);
        if (it != t.owner.end()) s->reqs[it->second].recs.push_back(r);
    }
    for (uint32_t slot : t.finishing) {
        Req &q = s->reqs[t.owner[slot]];
        q.t_done = t_now;
        q.checksum = t.h_sums[slot].checksum;
        q.prp_count = t.h_sums[slot].count;
        s->stats.completed++;
        s->stats.tested += t.h_sums[slot].count;
        s->done.push_back(q.id);
        s->active.erase(std::find(s->active.begin(), s->active.end(), q.id));
        s->free_slots.push_back(slot);
        /* every tick with its segments has finished (ticks retire in order) */
        arena_give(s, q.pre_off, q.blocks);
        s->stats.reserved_bytes -= q.blocks * kBlockBytes;
    }
    s->stats.ticks++;
    s->stats.segments += t.segments;
    t.busy = false;
    t.finishing.clear();
    t.owner.clear();
    s->cv.notify_all();
    return true;
}

void engine_main(ww_server *s)
{
    if (!hip_ok(s, hipSetDevice(s->device), "hipSetDevice")) return;
    const uint64_t T = s->cfg.tick_segments;
    for (;;) {
        /* admission: FIFO into free slots and free arena blocks (the HBM budget) */
        std::vector<uint32_t> cleared;
        std::vector<uint64_t> admitted;
        {
            std::unique_lock<std::mutex> g(s->mu);
            bool inflight = false;
            for (Tick &t : s->ticks) inflight |= t.busy;
            bool work = false;
            for (uint64_t id : s->active) work |= s->reqs[id].next < s->reqs[id].nseg;
            if (!work && !inflight && s->pending.empty()) {
                if (s->stop) return;
                s->cv.wait_for(g, std::chrono::milliseconds(2));
                continue;
            }
            while (!s->pending.empty() && !s->free_slots.empty()) {
                Req &q = s->reqs[s->pending.front()];
                const uint64_t need = blocks_for(s, q.nseg);
                if (!arena_take(s, need, &q.pre_off)) break; /* the head waits for HBM */
                q.blocks = need;
                s->pending.pop_front();
                q.slot = s->free_slots.back();
                s->free_slots.pop_back();
                q.t_admit = now_s();
                s->active.push_back(q.id);
                cleared.push_back(q.slot);
                if (s->nwide) admitted.push_back(q.id);
                s->stats.reserved_bytes += need * kBlockBytes;
                s->stats.peak_reserved_bytes =
                    std::max(s->stats.peak_reserved_bytes, s->stats.reserved_bytes);
                s->stats.peak_active = std::max<uint64_t>(s->stats.peak_active, s->active.size());
            }
        }
        for (uint32_t slot : cleared)
            if (!hip_ok(s, hipMemsetAsync(&s->d_sums[slot], 0, sizeof(SlotSum), s->stream), "memset"))
                return;
        /* each admitted request's sieve state, once, into its arena block: the one-span
           table is rewritten per request, in stream order behind the previous pre-sieve */
        for (uint64_t id : admitted) {
            uint64_t lo, hi, off, nseg;
            {
                std::lock_guard<std::mutex> g(s->mu);
                const Req &q = s->reqs[id];
                lo = q.lo;
                hi = q.hi;
                off = q.pre_off;
                nseg = q.nseg;
            }
            const int pgrid = (int)std::min<uint64_t>((nseg + s->presieve_chunk - 1) / s->presieve_chunk,
                                                      (uint64_t)(s->cus > 0 ? s->cus : 1));
            if (!hip_ok(s, ww::launch_set_span(lo, hi, s->d_adm_span, s->d_adm_prefix, s->stream), "span") ||
                !hip_ok(s, ww::launch_presieve(s->d_adm_span, s->d_adm_prefix, 1, s->d_wide, s->nwide,
                                               s->d_arena + off * ww::kPresieveWordsPerSeg,
                                               pgrid > 0 ? pgrid : 1, s->stream, s->presieve_chunk),
                           "k_presieve"))
                return;
        }

        Tick &t = s->ticks[s->ring_next];
        if (t.busy && !retire(s, t)) return;

        /* build the tick: next segments of active requests in admission order */
        uint32_t nspans = 0;
        uint64_t total = 0;
        {
            std::lock_guard<std::mutex> g(s->mu);
            t.h_prefix[0] = 0;
            for (uint64_t id : s->active) {
                if (total >= T || nspans >= s->cfg.slots) break;
                Req &q = s->reqs[id];
                uint64_t take = std::min(q.nseg - q.next, T - total);
                if (!take) continue;
                BatchSpan &sp = t.h_spans[nspans];
                sp.lo = q.lo;
                sp.hi = q.hi;
                sp.seg0 = q.next;
                sp.slot = q.slot;
                sp.pad = 0;
                sp.pre_seg = q.pre_off + q.next; /* its words in the request's block */
                q.next += take;
                total += take;
                t.h_prefix[++nspans] = total;
                t.owner[q.slot] = id;
                if (q.next == q.nseg) t.finishing.push_back(q.slot);
            }
        }
        if (!nspans) {
            /* nothing new to schedule: retire the oldest tick in flight */
            for (int k = 1; k <= kRing; ++k) {
                Tick &o = s->ticks[(s->ring_next + k) % kRing];
                if (o.busy) {
                    if (!retire(s, o)) return;
                    break;
                }
            }
            continue;
        }
        t.segments = total;
        bool ok = hip_ok(s, hipMemcpyAsync(t.d_spans, t.h_spans, nspans * sizeof(BatchSpan),
                                           hipMemcpyHostToDevice, s->stream), "spans") &&
                  hip_ok(s, hipMemcpyAsync(t.d_prefix, t.h_prefix, (nspans + 1) * sizeof(uint64_t),
                                           hipMemcpyHostToDevice, s->stream), "prefix") &&
                  hip_ok(s, ww::launch_batch(t.d_spans, t.d_prefix, nspans, t.d_state, s->d_sums, 0,
                                             t.d_res, kTickRecords, s->grid, s->cfg.claim, s->d_deep,
                                             s->ndeep, s->d_arena, s->stream, s->waves), "k_batch") &&
                  hip_ok(s, hipMemcpyAsync(t.h_state, t.d_state, sizeof(BatchState),
                                           hipMemcpyDeviceToHost, s->stream), "state") &&
                  hip_ok(s, hipMemcpyAsync(t.h_res, t.d_res, kTickCopy * sizeof(ww_result),
                                           hipMemcpyDeviceToHost, s->stream), "records") &&
                  hip_ok(s, hipMemcpyAsync(t.h_sums, s->d_sums, s->cfg.slots * sizeof(SlotSum),
                                           hipMemcpyDeviceToHost, s->stream), "sums") &&
                  hip_ok(s, hipEventRecord(t.done, s->stream), "event");
        if (!ok) return;
        t.busy = true;
        s->ring_next = (s->ring_next + 1) % kRing;
        /* retire whatever already finished, oldest first */
        for (int k = 0; k < kRing; ++k) {
            Tick &o = s->ticks[(s->ring_next + k) % kRing];
            if (o.busy && hipEventQuery(o.done) == hipSuccess && !retire(s, o)) return;
        }
    }
}

void free_tick(Tick &t)
{
    if (t.done) (void)hipEventDestroy(t.done);
    if (t.d_spans) (void)hipFree(t.d_spans);
    if (t.d_prefix) (void)hipFree(t.d_prefix);
    if (t.d_state) (void)hipFree(t.d_state);
    if (t.d_res) (void)hipFree(t.d_res);
    if (t.h_spans) (void)hipHostFree(t.h_spans);
    if (t.h_prefix) (void)hipHostFree(t.h_prefix);
    if (t.h_state) (void)hipHostFree(t.h_state);
    if (t.h_res) (void)hipHostFree(t.h_res);
    if (t.h_sums) (void)hipHostFree(t.h_sums);
}

void destroy(ww_server *s)
{
    if (s->engine.joinable()) {
        {
            std::lock_guard<std::mutex> g(s->mu);
            s->stop = true;
        }
        s->cv.notify_all();
        s->engine.join();
    }
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (Tick &t : s->ticks) free_tick(t);
    if (s->d_sums) (void)hipFree(s->d_sums);
    if (s->d_deep) (void)hipFree(s->d_deep);
    if (s->d_wide) (void)hipFree(s->d_wide);
    if (s->d_arena) (void)hipFree(s->d_arena);
    if (s->d_adm_span) (void)hipFree(s->d_adm_span);
    if (s->d_adm_prefix) (void)hipFree(s->d_adm_prefix);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

} // namespace

/* internal accessors of the context (ww_api.cpp) */
int ww_ctx_internals(ww_ctx *c, int *device, const ww::DeepPrime **deep, uint32_t *ndeep,
                     int *cus, int *tuned_blocks, uint32_t *tuned_claim,
                     const ww::WidePrime **wide, uint32_t *nwide, int *waves,
                     uint32_t *presieve_chunk);
int ww_set_error(int code, const char *fmt, ...);

extern "C" {

int ww_server_create(ww_ctx *ctx, const ww_server_config *cfg, ww_server **out)
{
    if (!ctx || !out) return ww_set_error(WW_E_ARG, "null argument");
    *out = nullptr;
    ww_server *s = new ww_server;
    s->ctx = ctx;
    s->cfg.slots = 4096;
    s->cfg.tick_segments = 1u << 17;
    s->cfg.hbm_budget = 64ull << 30;
    s->cfg.bytes_per_request = 0;
    s->cfg.claim = 0;
    s->cfg.blocks_per_cu = -1;
    if (cfg) {
        if (cfg->slots) s->cfg.slots = cfg->slots;
        if (cfg->tick_segments) s->cfg.tick_segments = cfg->tick_segments;
        if (cfg->hbm_budget) s->cfg.hbm_budget = cfg->hbm_budget;
        if (cfg->bytes_per_request) s->cfg.bytes_per_request = cfg->bytes_per_request;
        s->cfg.claim = cfg->claim;
        s->cfg.blocks_per_cu = cfg->blocks_per_cu;
    }
    int cus = 0, tuned_blocks = 0;
    uint32_t tuned_claim = 8;
    const ww::DeepPrime *ctx_deep = nullptr;
    const ww::WidePrime *ctx_wide = nullptr;
    int rc = ww_ctx_internals(ctx, &s->device, &ctx_deep, &s->ndeep, &cus, &tuned_blocks,
                              &tuned_claim, &ctx_wide, &s->nwide, &s->waves, &s->presieve_chunk);
    s->cus = cus;
    if (rc) {
        delete s;
        return rc;
    }
    /* ticks are short, so a wave claims one segment at a time unless told otherwise:
       longer claims leave long tails at the end of every tick (latency, not throughput) */
    (void)tuned_claim;
    if (!s->cfg.claim) s->cfg.claim = 1;
    int bpc = s->cfg.blocks_per_cu >= 0 ? s->cfg.blocks_per_cu : tuned_blocks;
    if (bpc <= 0) bpc = ww::batch_blocks_per_cu(s->waves);
    s->grid = cus * (bpc > 0 ? bpc : 4);
    {
        size_t free_b = 0, total_b = 0;
        if (hipSetDevice(s->device) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
            s->cfg.hbm_budget > free_b * 9 / 10)
            s->cfg.hbm_budget = free_b * 9 / 10; /* admission never exceeds free HBM */
    }
    s->arena_blocks = s->cfg.hbm_budget / kBlockBytes;
    s->extra_blocks = (s->cfg.bytes_per_request + kBlockBytes - 1) / kBlockBytes;
    if (s->arena_blocks) s->free_blocks.emplace(0, s->arena_blocks);
    for (uint32_t i = s->cfg.slots; i > 0; --i) s->free_slots.push_back(i - 1);

    hipError_t e = hipSetDevice(s->device);
    if (e == hipSuccess && s->ndeep) e = hipMalloc(&s->d_deep, s->ndeep * sizeof(ww::DeepPrime));
    if (e == hipSuccess && s->ndeep)
        e = hipMemcpy(s->d_deep, ctx_deep, s->ndeep * sizeof(ww::DeepPrime), hipMemcpyDeviceToDevice);
    if (e == hipSuccess && s->nwide) e = hipMalloc(&s->d_wide, s->nwide * sizeof(ww::WidePrime));
    if (e == hipSuccess && s->nwide)
        e = hipMemcpy(s->d_wide, ctx_wide, s->nwide * sizeof(ww::WidePrime), hipMemcpyDeviceToDevice);
    if (e == hipSuccess && s->nwide) e = hipMalloc(&s->d_arena, s->arena_blocks * kBlockBytes);
    if (e == hipSuccess && s->nwide) e = hipMalloc(&s->d_adm_span, sizeof(BatchSpan));
    if (e == hipSuccess && s->nwide) e = hipMalloc(&s->d_adm_prefix, 2 * sizeof(uint64_t));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&s->d_sums, s->cfg.slots * sizeof(SlotSum));
    if (e == hipSuccess) e = hipMemsetAsync(s->d_sums, 0, s->cfg.slots * sizeof(SlotSum), s->stream);
    for (Tick &t : s->ticks) {
        if (e == hipSuccess) e = hipMalloc(&t.d_spans, s->cfg.slots * sizeof(BatchSpan));
        if (e == hipSuccess) e = hipMalloc(&t.d_prefix, (s->cfg.slots + 1) * sizeof(uint64_t));
        if (e == hipSuccess) e = hipMalloc(&t.d_state, sizeof(BatchState));
        if (e == hipSuccess) e = hipMalloc(&t.d_res, kTickRecords * sizeof(ww_result));
        if (e == hipSuccess) e = hipHostMalloc(&t.h_spans, s->cfg.slots * sizeof(BatchSpan), 0);
        if (e == hipSuccess) e = hipHostMalloc(&t.h_prefix, (s->cfg.slots + 1) * sizeof(uint64_t), 0);
        if (e == hipSuccess) e = hipHostMalloc(&t.h_state, sizeof(BatchState), 0);
        if (e == hipSuccess) e = hipHostMalloc(&t.h_res, kTickRecords * sizeof(ww_result), 0);
        if (e == hipSuccess) e = hipHostMalloc(&t.h_sums, s->cfg.slots * sizeof(SlotSum), 0);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&t.done, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) {
        destroy(s);
        return ww_set_error(WW_E_HIP, "ww_server_create: %s", hipGetErrorString(e));
    }
    s->stats.t_start = now_s();
    s->engine = std::thread(engine_main, s);
    *out = s;
    return WW_OK;
}

int ww_server_submit(ww_server *s, uint64_t lo, uint64_t hi, uint64_t *id)
{
    if (!s) return ww_set_error(WW_E_ARG, "null server");
    if (lo < WW_MIN_START || hi <= lo) return ww_set_error(WW_E_ARG, "bad range");
    std::lock_guard<std::mutex> g(s->mu);
    if (s->error) return ww_set_error(s->error, "%s", s->errmsg);
    Req q;
    q.id = s->next_id++;
    q.lo = lo;
    q.hi = hi;
    q.nseg = ww::segments_of(lo, hi);
    if (blocks_for(s, q.nseg) > s->arena_blocks)
        return ww_set_error(WW_E_CAPACITY,
                            "request [%llu, %llu) needs %llu bytes of resident sieve state, above "
                            "the HBM budget of %llu bytes",
                            (unsigned long long)lo, (unsigned long long)hi,
                            (unsigned long long)(blocks_for(s, q.nseg) * kBlockBytes),
                            (unsigned long long)(s->arena_blocks * kBlockBytes));
    q.t_submit = now_s();
    s->stats.submitted++;
    if (q.nseg == 0) { /* no odd number in range: complete at once */
        q.t_admit = q.t_done = q.t_submit;
        s->stats.completed++;
        s->done.push_back(q.id);
    } else {
        s->pending.push_back(q.id);
    }
    if (id) *id = q.id;
    s->reqs.emplace(q.id, std::move(q));
    s->cv.notify_all();
    return WW_OK;
}

int ww_server_poll(ww_server *s, ww_request_info *out, uint32_t cap, uint32_t *n, int wait_ms)
{
    if (!s || !n) return ww_set_error(WW_E_ARG, "null argument");
    std::unique_lock<std::mutex> g(s->mu);
    if (wait_ms > 0 && s->done.empty() && !s->error)
        s->cv.wait_for(g, std::chrono::milliseconds(wait_ms),
                       [s] { return !s->done.empty() || s->error; });
    if (s->error) return ww_set_error(s->error, "%s", s->errmsg);
    uint32_t k = 0;
    while (k < cap && !s->done.empty()) {
        uint64_t id = s->done.front();
        s->done.pop_front();
        Req &q = s->reqs[id];
        ww_request_info &r = out[k++];
        r.id = q.id;
        r.lo = q.lo;
        r.hi = q.hi;
        r.checksum = q.checksum;
        r.prp_count = q.prp_count;
        r.t_submit = q.t_submit - s->stats.t_start;
        r.t_admit = q.t_admit - s->stats.t_start;
        r.t_done = q.t_done - s->stats.t_start;
        r.nrecords = (uint32_t)q.recs.size();
        r.pad = 0;
        if (!q.recs.empty()) s->taken.emplace(id, std::move(q.recs));
        s->reqs.erase(id);
    }
    *n = k;
    return WW_OK;
}

int ww_server_records(ww_server *s, uint64_t id, ww_result *out, uint32_t cap, uint32_t *n)
{
    if (!s || !n) return ww_set_error(WW_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->taken.find(id);
    uint32_t k = 0;
    if (it != s->taken.end()) {
        k = (uint32_t)std::min<size_t>(cap, it->second.size());
        if (k && !out) return ww_set_error(WW_E_ARG, "null out");
        std::copy(it->second.begin(), it->second.begin() + k, out);
        s->taken.erase(it);
    }
    *n = k;
    return WW_OK;
}

int ww_server_get_stats(ww_server *s, ww_server_stats *out)
{
    if (!s || !out) return ww_set_error(WW_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->stats;
    out->pending = s->pending.size();
    out->active = s->active.size();
    return WW_OK;
}

int ww_server_destroy(ww_server *s)
{
    if (s) destroy(s);
    return WW_OK;
}

} // extern "C"
