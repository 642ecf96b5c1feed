/*
 * ww_batch.hip -- the persistent form: one launch searches many ranges.
 *
 * Work unit: a segment of 8192 consecutive integers (64 lanes x one 64-bit word of odd
 * numbers each) of one range. The ranges' segments are laid end to end ("spans", with a
 * prefix sum of their segment counts) and every wave claims the next segment from one
 * global cursor (work stealing), so uneven ranges balance by construction. Per segment a
 * wave sieves by 3..113 in registers (and optionally by 127..B in LDS, see deep_sieve),
 * compacts candidates into LDS, runs the base-2
 * strong PRP 64 candidates at a time and queues the survivors (p, range slot) in LDS;
 * whenever 64 are queued it runs the Wieferich + Wall-Sun-Sun tests on them, one per
 * lane, so the test lanes stay full across segment and range boundaries. Checksums and
 * tested counts accumulate per range slot (a wave keeps a running sum while its slot
 * does not change); result records carry the slot in their pad field.
 *
 * Replaces the reference's per-chunk launch sequence (main.cpp:1683-1741) for many
 * ranges at once; the arithmetic and classification are those of the plain kernels.
 */
#include "ww_batch.h"
#include "ww_device.h"

namespace ww {

namespace {

constexpr int kQueueCap = 128; /* < 64 left over + <= 64 appended per PRP round */
/* candidates awaiting the strong PRP test outside the candidate buffer (p, range slot):
   fewer than 64 between rounds, plus one segment's when they fit */
constexpr u32 kPspCap = 96;
/* candidates of one segment (at most kWaveCandCap) behind fewer than 64 leftovers */
constexpr int kBatchCandCap = kWaveCandCap + 64;
static_assert(kBatchCandCap - 256 >= 64, "rescue bitmap must not reach the leftovers");

__device__ __forceinline__ void wave_sync_lds()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WaveAcc {
    u32 slot;
    u64 cks, cnt;
};

__device__ __forceinline__ void acc_flush(WaveAcc &a, SlotSum *sums, u32 lane)
{
    if (a.slot != 0xffffffffu && lane == 0 && a.cnt) {
        atomicAdd((unsigned long long *)&sums[a.slot].checksum, (unsigned long long)a.cks);
        atomicAdd((unsigned long long *)&sums[a.slot].count, (unsigned long long)a.cnt);
    }
    a.cks = 0;
    a.cnt = 0;
}

/* Tests the first n (1..64) queued entries, one per lane, and drops them from the queue. */
__device__ __forceinline__ void tests_round(u32 n, u32 &qn, u64 *wq, u32 *ws, u32 lane,
                                            WaveAcc &acc, SlotSum *sums, ww_result *res,
                                            u32 res_cap, u32 *rcount)
{
    const bool act = lane < n;
    const u64 p = act ? wq[lane] : 0;
    const u32 slot = act ? ws[lane] : 0;
    const u32 rest = qn - n;
    u64 pr = 0;
    u32 sr = 0;
    if (lane < rest) {
        pr = wq[n + lane];
        sr = ws[n + lane];
    }
    wave_sync_lds();
    if (lane < rest) {
        wq[lane] = pr;
        ws[lane] = sr;
    }
    wave_sync_lds();
    qn = rest;

    const u64 v = test_prime_wave(p, act, res, res_cap, rcount, (int32_t)slot);

    /* per-slot sums: one wave reduction per distinct slot in the round (entries of one
       segment are contiguous, so a round holds one slot or a few) */
    u64 todo = __ballot(act);
    while (todo) {
        const u32 si = uniform32(__shfl(slot, (int)__builtin_ctzll(todo), 64));
        const bool mine = act && slot == si;
        const u64 m = __ballot(mine);
        u64 sum = mine ? v : 0;
        for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
        sum = uniform64(sum);
        if (si != acc.slot) {
            acc_flush(acc, sums, lane);
            acc.slot = si;
        }
        acc.cks += sum;
        acc.cnt += (u64)__popcll(m);
        todo &= ~m;
    }
}

/* Wave compaction of a segment's candidates as compact_by_class does it, split in two:
   the candidates in `need` (those the strong PRP test must decide) and the others, known
   primes, which go to the candidate buffer as one group tagged with kKnown, so that PRP
   rounds over them skip the test. The candidates to test go to the PRP queue pq/ps (as
   p = G + 2 i with the range slot) when they all fit behind its pn entries; otherwise they
   go to the candidate buffer first, grouped by the trailing zeros of n - 1 in the order
   `desc` gives. The five counts share one 64-bit scan in 12-bit fields (a segment holds at
   most kWaveCandCap < 4096 candidates). Returns the number of entries written to the
   buffer, wave-uniform, and advances pn. Called by all 64 lanes. */
constexpr u32 kKnown = 1u << 13;

__device__ __forceinline__ u32 compact_split(u64 bits, u64 need, u64 G, u32 lane, uint16_t *wc,
                                             u32 base, u32 tag, bool desc, u64 *pq, u32 *ps,
                                             u32 &pn, u32 slot)
{
    const u32 g8 = (u32)((G - 1) >> 1) & 7u;
    u64 cls[4];
    cls[0] = 0x5555555555555555ull << ((1u - g8) & 1u); /* g + i odd: t = 1 */
    cls[1] = 0x1111111111111111ull << ((2u - g8) & 3u); /* = 2 mod 4: t = 2 */
    cls[2] = 0x0101010101010101ull << ((4u - g8) & 7u); /* = 4 mod 8: t = 3 */
    cls[3] = 0x0101010101010101ull << ((0u - g8) & 7u); /* = 0 mod 8: t >= 4 */
    const u64 test = bits & need, known = bits & ~need;
    u64 cnt = (u64)__popcll(known) << 48;
#pragma unroll
    for (int k = 0; k < 4; ++k) cnt |= (u64)__popcll(test & cls[k]) << (12 * k);
    u64 incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
        u64 v = __shfl_up(incl, d, 64);
        if (lane >= (u32)d) incl += v;
    }
    const u64 tot = uniform64(__shfl(incl, 63, 64));
    const u64 excl = incl - cnt;
    const u32 ptot = (u32)((tot & 0xfffu) + ((tot >> 12) & 0xfffu) + ((tot >> 24) & 0xfffu) +
                           ((tot >> 36) & 0xfffu));
    u32 at = base;
    if (pn + ptot <= kPspCap) {
        u32 pos = pn + (u32)((excl & 0xfffu) + ((excl >> 12) & 0xfffu) + ((excl >> 24) & 0xfffu) +
                             ((excl >> 36) & 0xfffu));
        u64 b = test;
        while (b) {
            pq[pos] = G + 2ull * (lane * 64 + (u32)__builtin_ctzll(b));
            ps[pos] = slot;
            ++pos;
            b &= b - 1;
        }
        pn += ptot;
    } else {
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int k = desc ? 3 - kk : kk;
            u32 pos = at + (u32)((excl >> (12 * k)) & 0xfffu);
            u64 b = test & cls[k];
            while (b) {
                wc[pos++] = (uint16_t)(tag | (lane * 64 + (u32)__builtin_ctzll(b)));
                b &= b - 1;
            }
            at += (u32)((tot >> (12 * k)) & 0xfffu);
        }
    }
    u32 pos = at + (u32)((excl >> 48) & 0xfffu);
    u64 b = known;
    while (b) {
        wc[pos++] = (uint16_t)(kKnown | tag | (lane * 64 + (u32)__builtin_ctzll(b)));
        b &= b - 1;
    }
    at += (u32)((tot >> 48) & 0xfffu);
    return at - base;
}

/* Waves per SIMD each shape is compiled for (96 VGPRs at 5, 128 at 4). Eight-wave
   workgroups (54 KB of LDS each) fit two per CU, four waves per SIMD. */
template <int W> struct BatchShape {
    static constexpr int kWavesPerEU = W == 8 ? 4 : 5;
};

template <int W>
__global__ __launch_bounds__(64 * W)
__attribute__((amdgpu_waves_per_eu(BatchShape<W>::kWavesPerEU))) void k_batch(const BatchSpan *spans,
                                                             const u64 *prefix, u32 nspans,
                                                             BatchState *st, SlotSum *sums,
                                                             ww_result *res, u32 res_cap,
                                                             u32 claim, const DeepPrime *deep,
                                                             u32 ndeep, const u64 *pre,
                                                             const u64 *prer)
{
    __shared__ __align__(16) uint16_t cand[W][kBatchCandCap];
    __shared__ u64 deep_bm[W][64];
    __shared__ u64 qp[W][kQueueCap];
    __shared__ u32 qs[W][kQueueCap];
    __shared__ u64 pp[W][kPspCap];
    __shared__ u32 psl[W][kPspCap];

    const u32 lane = threadIdx.x & 63;
    const u32 wave = threadIdx.x >> 6;
    uint16_t *wc = cand[wave];
    u64 *wq = qp[wave];
    u32 *ws = qs[wave];
    u64 *pq = pp[wave];
    u32 *ps = psl[wave];
    const u64 total = prefix[nspans];

    u32 qn = 0, pn = 0;
    WaveAcc acc;
    acc.slot = 0xffffffffu;
    acc.cks = acc.cnt = 0;

    /* One wave-uniform state machine with a single site per phase (keeps each phase's
       registers to itself): drain a full test round, else run the next PRP round, else
       claim and sieve the next segment.
       Candidates are 12-bit odd-number positions tagged with the segment they belong to:
       bit 12 clear = the older segment (base g_old, range slot s_old), set = the newest
       (g_new, s_new). When fewer than 64 candidates of the newest segment remain, they move
       to the front as "older" ones and the next segment's candidates are appended behind
       them, so PRP rounds stay full across segment boundaries. */
    bool done = false;
    u32 j0 = 0, ncand = 0, kb = 0; /* kb: first entry of the newest segment */
    u32 s_old = 0, s_new = 0;
    u64 g_old = 0, g_new = 0;
    u64 seg_next = 0, seg_end = 0; /* segments claimed but not yet sieved */
    u64 seen = 0;                  /* cursor after this wave's last claim */
    bool desc = true;              /* group order of the newest segment's candidates */
    for (;;) {
        if (qn >= 64 || (done && qn > 0)) {
            tests_round(qn >= 64 ? 64 : qn, qn, wq, ws, lane, acc, sums, res, res_cap,
                        &st->resultcount);
            continue;
        }
        const u32 rem = j0 < ncand ? ncand - j0 : 0;
        const bool from_buf = rem >= 64 || (rem > 0 && (done || j0 < kb));
        if (from_buf || pn >= 64 || (done && pn > 0)) {
            /* 64 candidates per round, from the candidate buffer or else from the PRP
               queue; known primes pass, the rest take the strong PRP test (skipped when
               none is in the round); survivors join the test queue */
            u64 n;
            u32 sl;
            bool in, test;
            if (from_buf) {
                const u32 j = j0 + lane;
                in = j < ncand;
                const u32 e = wc[in ? j : j0];
                const bool newest = (e >> 12) & 1;
                n = (newest ? g_new : g_old) + 2ull * (e & 0xfffu);
                sl = newest ? s_new : s_old;
                test = in && !(e & kKnown);
                j0 += 64;
            } else {
                const u32 nr = pn >= 64 ? 64 : pn, rest = pn - nr;
                in = lane < nr;
                n = in ? pq[lane] : 0;
                sl = in ? ps[lane] : 0;
                test = in;
                u64 pr = 0;
                u32 sr = 0;
                if (lane < rest) {
                    pr = pq[nr + lane];
                    sr = ps[nr + lane];
                }
                wave_sync_lds();
                if (lane < rest) {
                    pq[lane] = pr;
                    ps[lane] = sr;
                }
                pn = rest;
            }
            bool prp = in;
            if (__ballot(test)) {
                const bool r = spsp2_wave(n, test); /* all lanes: it takes wave-wide bounds */
                prp = test ? r : prp;
            }
            const u64 bal = __ballot(prp);
            if (prp) {
                const u32 at = qn + (u32)__popcll(bal & ((1ull << lane) - 1));
                wq[at] = n;
                ws[at] = sl;
            }
            qn += (u32)__popcll(bal);
            wave_sync_lds();
            continue;
        }
        if (done) break;
        if (seg_next >= seg_end) {
            /* claim consecutive segments with one atomic (runs of one range): up to `claim`
               of them, tapering to one as the work runs out (remaining segments over four
               per wave of the grid, as last seen by this wave), so the launch ends with
               short claims and no long tail */
            const u64 left = total > seen ? total - seen : 0;
            u64 k = left / ((u64)gridDim.x * W * 4);
            k = k < 1 ? 1 : (k > claim ? claim : k);
            u64 c = 0;
            if (lane == 0) c = atomicAdd((unsigned long long *)&st->cursor, (unsigned long long)k);
            seg_next = uniform64(__shfl(c, 0, 64));
            seen = seg_next + k;
            seg_end = seg_next + k < total ? seg_next + k : total;
            if (seg_next >= total) {
                done = true;
                continue;
            }
        }
        const u64 s = seg_next++;
        /* span holding segment s: largest a with prefix[a] <= s */
        u32 a = 0, b = nspans;
        while (b - a > 1) {
            u32 m = (a + b) >> 1;
            if (prefix[m] <= s) a = m;
            else b = m;
        }
        const BatchSpan sp = spans[a];

        /* the rem (< 64) leftovers, all of the newest segment, become the older ones at
           the front of the buffer (rem <= j0 unless j0 = 0, so the moves never overlap) */
        {
            u32 e = 0;
            if (lane < rem) e = wc[j0 + lane] & ~(1u << 12);
            wave_sync_lds();
            if (lane < rem) wc[lane] = (uint16_t)e;
            wave_sync_lds();
        }
        g_old = g_new;
        s_old = s_new;
        s_new = uniform32(sp.slot);
        g_new = uniform64((sp.lo | 1) + (sp.seg0 + (s - prefix[a])) * kIntsPerWave);

        /* sieve, then compact candidate positions into LDS behind the leftovers */
        u64 bits = sieve_word(g_new, lane * (u32)kIntsPerWord, sp.hi);
        u64 need = ~0ull; /* candidates the strong PRP test must decide */
        if (pre) {
            const u64 wi = (sp.pre_seg + (s - prefix[a])) * kPresieveWordsPerSeg + lane;
            bits &= ~pre[wi];
            if (prer) need = prer[wi];
        }
        if (ndeep) {
            /* the end of the candidate buffer is free until the compaction below: it holds
               the deep sieve's pseudoprime classes meanwhile (the leftovers sit at the
               front) */
            u64 *bm = deep_bm[wave];
            u64 *rm = reinterpret_cast<u64 *>(wc + kBatchCandCap - 256);
            bm[lane] = 0;
            rm[lane] = 0;
            wave_sync_lds();
            u64 classes;
            bits &= ~deep_sieve(bm, rm, g_new, deep, ndeep, lane, &classes);
            need |= classes;
        }
        /* compacted behind the leftovers: the candidates to test to the PRP queue when they
           fit, else here grouped by the trailing zeros of n - 1, in the opposite order to
           the previous segment's, whose last group the leftovers are; then the known
           primes */
        kb = rem;
        j0 = 0;
        desc = !desc;
        ncand = rem + compact_split(bits, need, g_new, lane, wc, rem, 1u << 12, desc, pq, ps, pn, s_new);
        wave_sync_lds();
    }
    acc_flush(acc, sums, lane);
}

__global__ void k_batch_clear(BatchState *st, SlotSum *sums, u32 nslots)
{
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        st->cursor = 0;
        st->resultcount = 0;
        st->reserved = 0;
    }
    if (i < nslots) {
        sums[i].checksum = 0;
        sums[i].count = 0;
    }
}

/* One workgroup: counters by thread 0, then the records copied in parallel. */
__global__ void k_fold(const BatchState *bst, const SlotSum *sum, const ww_result *bres, u32 bcap,
                       DevState *ds, ww_result *res, u32 cap)
{
    __shared__ u32 base, n;
    if (threadIdx.x == 0) {
        const u32 made = bst->resultcount;
        n = made < bcap ? made : bcap;
        base = ds->resultcount;
        ds->resultcount = base + made;
        ds->checksum += sum->checksum;
        ds->totalprimecount += sum->count;
    }
    __syncthreads();
    for (u32 i = threadIdx.x; i < n; i += blockDim.x) {
        const u32 at = base + i;
        if (at < cap) {
            ww_result r = bres[i];
            r.pad = 0;
            res[at] = r;
        }
    }
}

} // namespace

hipError_t launch_fold(const BatchState *bst, const SlotSum *sum, const ww_result *bres,
                       u32 bcap, DevState *ds, ww_result *res, u32 cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, s, bst, sum, bres, bcap, ds, res, cap);
    return hipGetLastError();
}

bool batch_waves_ok(int waves) { return waves == 2 || waves == 4 || waves == 8; }

int batch_blocks_per_cu(int waves)
{
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
    if (waves == 2) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<2>, 128, 0);
    if (waves == 4) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<4>, 256, 0);
    if (waves == 8) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_batch<8>, 512, 0);
    return e == hipSuccess ? nb : 0;
}

hipError_t launch_batch(const BatchSpan *spans, const u64 *prefix, u32 nspans, BatchState *st,
                        SlotSum *sums, u32 nslots, ww_result *res, u32 res_cap, int grid,
                        u32 claim, const DeepPrime *deep, u32 ndeep, const u64 *pre,
                        hipStream_t s, int waves, const u64 *prer)
{
    if (!batch_waves_ok(waves)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_batch_clear, dim3(nslots > 256 ? (nslots + 255) / 256 : 1), dim3(256), 0, s,
                       st, sums, nslots);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const u32 c = claim < 1 ? 1u : claim;
    if (waves == 2)
        hipLaunchKernelGGL(k_batch<2>, dim3(grid), dim3(128), 0, s, spans, prefix, nspans, st, sums,
                           res, res_cap, c, deep, ndeep, pre, pre ? prer : nullptr);
    else if (waves == 8)
        hipLaunchKernelGGL(k_batch<8>, dim3(grid), dim3(512), 0, s, spans, prefix, nspans, st, sums,
                           res, res_cap, c, deep, ndeep, pre, pre ? prer : nullptr);
    else
        hipLaunchKernelGGL(k_batch<4>, dim3(grid), dim3(256), 0, s, spans, prefix, nspans, st, sums,
                           res, res_cap, c, deep, ndeep, pre, pre ? prer : nullptr);
    return hipGetLastError();
}

} // namespace ww
