/*
 * ww_kernels.hip -- gfx950 kernels of the search path.
 *
 *   k_sieve_prp   segmented candidate generation (odd-only bitsieve by 3..113 in
 *                 registers, optional deep sieve by 127..B in LDS with exact rescue of
 *                 base-2 pseudoprimes), wave-local compaction into LDS, base-2 strong PRP on the
 *                 compacted candidates, in-place compaction of the survivors and one
 *                 global append per workgroup. With the bucket sieve's needs-PRP words,
 *                 candidates outside them are proven prime and join the list without the
 *                 test. Replaces getsegprimes.cl:228-301.
 *   k_ww_tests    Wieferich + Wall-Sun-Sun on every list entry, one entry per lane,
 *                 wave-reduced checksum, result records. Replaces wieferich.cl:196-328
 *                 and wallsunsun.cl:221-354.
 *   k_residues, k_prp_flags
 *                 per-value access to the same device arithmetic for verification.
 */
#include "ww_device.h"

namespace ww {

__global__ __launch_bounds__(kSieveThreads) void k_sieve_prp(u64 base, u64 hi, u64 *list,
                                                              u64 cap, DevState *st,
                                                              const DeepPrime *deep, u32 ndeep,
                                                              const u64 *pre, const u64 *prer)
{
    __shared__ __align__(16) uint16_t cand[kSieveWaves * kWaveCandCap];
    __shared__ u64 deep_bm[kSieveWaves][64];
    __shared__ u32 s_cnt[kSieveWaves];
    __shared__ u32 s_off;

    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63;
    const u32 wave = tid >> 6;
    const u64 block_base = base + (u64)blockIdx.x * kIntsPerBlock;
    const u64 wave_base = block_base + (u64)wave * kIntsPerWave;
    const u32 lane_pos = tid * (u32)kIntsPerWord; /* offset of w0 from block_base */

    /* ---- sieve this lane's 64 odd numbers ---- */
    u64 bits = sieve_word(block_base, lane_pos, hi);
    /* wide pre-sieve of this wave's segment (segments count from base); with the bucket
       sieve also the candidates the strong PRP test must decide (the others are proven
       prime: every prime factor up to sqrt(hi) would have dropped or flagged them) */
    u64 need = ~0ull;
    if (pre && wave_base < hi) {
        const u64 w = ((u64)blockIdx.x * kSieveWaves + wave) * 64 + lane;
        bits &= ~pre[w];
        if (prer) need = prer[w];
    }
    u64 classes = 0;
    if (ndeep) {
        /* the candidate buffer is free until the compaction below: it holds the deep
           sieve's pseudoprime classes meanwhile */
        u64 *bm = deep_bm[wave];
        u64 *rm = reinterpret_cast<u64 *>(cand + wave * kWaveCandCap);
        bm[lane] = 0;
        rm[lane] = 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        bits &= ~deep_sieve(bm, rm, wave_base, deep, ndeep, lane, &classes);
    }
    need |= classes;
    const u64 known = bits & ~need;
    bits &= need;

    /* ---- wave-local compaction of candidate positions into LDS, grouped by the
       trailing zeros of n - 1 ---- */
    uint16_t *wc = cand + wave * kWaveCandCap;
    const u32 ncand = compact_by_class(bits, wave_base, lane, wc, 0, 0);
    /* the known primes' count before this lane (a wave prefix sum) and in all */
    u32 kpre = (u32)__popcll(known);
    for (int d = 1; d < 64; d <<= 1) {
        const u32 t = __shfl_up(kpre, d, 64);
        if (lane >= (u32)d) kpre += t;
    }
    const u32 nknown = __shfl(kpre, 63, 64);
    kpre -= (u32)__popcll(known);
    __syncthreads();

    /* ---- strong PRP over the compacted candidates, survivors compacted in place ---- */
    u32 nout = 0;
    for (u32 j0 = 0; j0 < ncand; j0 += 64) {
        u32 j = j0 + lane;
        bool act = j < ncand;
        u32 idx = act ? wc[j] : 0;
        const bool prp = spsp2_wave(wave_base + 2ull * idx, act);
        u64 bal = __ballot(prp);
        if (prp) {
            u32 below = __popcll(bal & ((1ull << lane) - 1));
            wc[nout + below] = (uint16_t)idx; /* nout + below <= j: already read */
        }
        nout += __popcll(bal);
    }
    /* the known primes after the PRP survivors (together at most the wave's candidates) */
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (u64 b = known; b; b &= b - 1)
        wc[nout + kpre++] = (uint16_t)(lane * 64 + (u32)__builtin_ctzll(b));
    nout += nknown;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

    /* ---- one global append per workgroup ---- */
    if (lane == 0) s_cnt[wave] = nout;
    __syncthreads();
    if (tid == 0) {
        u32 total = 0;
        for (int w = 0; w < kSieveWaves; ++w) total += s_cnt[w];
        s_off = total ? atomicAdd(&st->primecount, total) : 0;
    }
    __syncthreads();
    u64 out0 = s_off;
    for (int w = 0; w < (int)wave; ++w) out0 += s_cnt[w];
    for (u32 k = lane; k < nout; k += 64) {
        u64 at = out0 + k;
        if (at < cap) list[at] = wave_base + 2ull * wc[k];
    }
}

/* ---------------------------------------------------------------------------
 * Wieferich + Wall-Sun-Sun over the list
 * ------------------------------------------------------------------------- */

#ifndef WW_TESTS_WAVES_PER_EU
/* 80 VGPRs and 18 spilled; 4% faster than 4 waves at 108. 8 waves (64 VGPRs, 36 spilled)
   are 1.4% faster still, but their spill stores leave the L2: 3.4 GB written per 2^32 chunk
   at 2^63 instead of 18 MB */
#define WW_TESTS_WAVES_PER_EU 6
#endif

__global__ __launch_bounds__(kTestThreads)
__attribute__((amdgpu_waves_per_eu(WW_TESTS_WAVES_PER_EU))) void k_ww_tests(const u64 *list,
                                                                          u64 cap, DevState *st,
                                                                          ww_result *res,
                                                                          u32 res_cap)
{
    const u32 count = st->primecount;
    const u64 n = count < cap ? count : cap;
    u64 sum = 0;
    const u32 lane = threadIdx.x & 63;
    /* grid-stride over wave-sized pieces, so that a wave's lanes run the loop together */
    for (u64 base = (u64)blockIdx.x * kTestThreads + (threadIdx.x - lane); base < n;
         base += (u64)gridDim.x * kTestThreads) {
        const u64 i = base + lane;
        const bool act = i < n;
        sum += test_prime_wave(act ? list[i] : 0, act, res, res_cap, &st->resultcount, 0);
    }
    /* wave reduction, one 64-bit atomic per wave */
    for (int d = 32; d > 0; d >>= 1) sum += __shfl_xor(sum, d, 64);
    if ((threadIdx.x & 63) == 0 && sum) atomicAdd((unsigned long long *)&st->checksum, sum);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->totalprimecount += count; /* wieferich.cl:317-326 */
        if (count > st->primecount_max) st->primecount_max = count;
    }
}

__global__ void k_residues(const u64 *pin, u64 n, u64 *out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    u64 p = pin[i];
    u64 q = inv64(p);
    m2 one = m2_one(p);
    m2 s = wieferich_residue(p, q, one);
    m2 f = wallsunsun_residue(p, q, one);
    out[4 * i + 0] = s.d0;
    out[4 * i + 1] = s.d1;
    out[4 * i + 2] = f.d0;
    out[4 * i + 3] = f.d1;
}

__global__ void k_prp_flags(const u64 *nin, u64 count, uint8_t *out)
{
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    u64 n = nin[i];
    out[i] = spsp2(n, inv64(n), mont_one(n)) ? 1 : 0;
}

/* ---------------------------------------------------------------------------
 * Launchers
 * ------------------------------------------------------------------------- */

int tests_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_ww_tests, kTestThreads, 0) != hipSuccess)
        return 0;
    return nb;
}

hipError_t launch_sieve_prp(u64 lo, u64 hi, u64 *list, u64 cap, DevState *st,
                            const DeepPrime *deep, u32 ndeep, const u64 *pre, hipStream_t s,
                            const u64 *prer)
{
    u64 base = lo | 1;
    if (base >= hi) return hipSuccess;
    u64 span = hi - base;
    u64 blocks = (span + kIntsPerBlock - 1) / kIntsPerBlock;
    hipLaunchKernelGGL(k_sieve_prp, dim3((unsigned)blocks), dim3(kSieveThreads), 0, s, base, hi,
                       list, cap, st, deep, ndeep, pre, pre ? prer : nullptr);
    return hipGetLastError();
}

hipError_t launch_ww_tests(const u64 *list, u64 cap, DevState *st, ww_result *res, u32 res_cap,
                           int grid, hipStream_t s)
{
    hipLaunchKernelGGL(k_ww_tests, dim3(grid), dim3(kTestThreads), 0, s, list, cap, st, res,
                       res_cap);
    return hipGetLastError();
}

hipError_t launch_residues(const u64 *p, u64 n, u64 *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_residues, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, out);
    return hipGetLastError();
}

hipError_t launch_prp_flags(const u64 *n, u64 count, uint8_t *out, hipStream_t s)
{
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_prp_flags, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, n,
                       count, out);
    return hipGetLastError();
}

} // namespace ww
