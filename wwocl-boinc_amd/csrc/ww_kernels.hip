/*
 * ww_kernels.hip -- gfx950 kernels of the search path.
 *
 *   k_sieve_prp   segmented candidate generation (odd-only bitsieve by 3..113 in
 *                 registers), wave-local compaction into LDS, base-2 strong PRP on the
 *                 compacted candidates, in-place compaction of the survivors and one
 *                 global append per workgroup. Replaces getsegprimes.cl:228-301.
 *   k_ww_tests    Wieferich + Wall-Sun-Sun on every list entry, one entry per lane,
 *                 wave-reduced checksum, result records. Replaces wieferich.cl:196-328
 *                 and wallsunsun.cl:221-354.
 *   k_residues, k_prp_flags
 *                 per-value access to the same device arithmetic for verification.
 */
#include "ww_kernels.h"
#include "ww_arith.h"

namespace ww {

/* ---------------------------------------------------------------------------
 * Sieve by the odd primes 3..113: the candidate set of the reference's wheel +
 * bit tables (getsegprimes.cl:167-278) and of its CPU loop (main.cpp:653-682).
 * ------------------------------------------------------------------------- */

/* Bits 0, L, 2L, ... of a 64-bit word: one lane's word holds odd numbers
   w0, w0+2, ..., w0+126, so multiples of L recur every L bits. */
__device__ constexpr u64 stride_pattern(u32 L)
{
    u64 m = 0;
    for (u32 i = 0; i < 64; i += L) m |= 1ull << i;
    return m;
}

/* Marks the bits of `comp` whose odd number is divisible by L. base_mod = w0_block mod
   L, off = (position of this lane's word relative to the block) mod L is folded in. */
template <u32 L>
__device__ __forceinline__ void sieve_prime(u64 &comp, u32 block_mod, u32 lane_off)
{
    constexpr u32 inv2 = (L + 1) / 2; /* 2^-1 mod L */
    constexpr u64 pat = stride_pattern(L);
    u32 r = (block_mod + lane_off % L) % L;  /* w0 mod L */
    u32 i0 = ((L - r) * inv2) % L;          /* first i with L | w0 + 2i */
    if (L < 64) {
        comp |= pat << i0;
    } else {
        comp |= (i0 < 64) ? (1ull << i0) : 0ull;
    }
}

#define WW_SIEVE_PRIMES(X)                                                            \
    X(3) X(5) X(7) X(11) X(13) X(17) X(19) X(23) X(29) X(31) X(37) X(41) X(43) X(47) \
    X(53) X(59) X(61) X(67) X(71) X(73) X(79) X(83) X(89) X(97) X(101) X(103) X(107) \
    X(109) X(113)

__global__ __launch_bounds__(kSieveThreads) void k_sieve_prp(u64 base, u64 hi, u64 *list,
                                                              u64 cap, DevState *st)
{
    __shared__ uint16_t cand[kSieveWaves * kWaveCandCap];
    __shared__ u32 s_cnt[kSieveWaves];
    __shared__ u32 s_off;

    const u32 tid = threadIdx.x;
    const u32 lane = tid & 63;
    const u32 wave = tid >> 6;
    const u64 block_base = base + (u64)blockIdx.x * kIntsPerBlock;
    const u64 wave_base = block_base + (u64)wave * kIntsPerWave;
    const u64 w0 = wave_base + (u64)lane * kIntsPerWord;
    const u32 lane_pos = tid * (u32)kIntsPerWord; /* offset of w0 from block_base */

    /* ---- sieve this lane's 64 odd numbers ---- */
    u64 comp = 0;
#define WW_SIEVE_ONE(L) sieve_prime<L>(comp, (u32)(block_base % L), lane_pos);
    WW_SIEVE_PRIMES(WW_SIEVE_ONE)
#undef WW_SIEVE_ONE

    /* keep odd n = w0 + 2i with n < hi (n >= base holds by construction); words past
       2^64 wrap below block_base and are dropped */
    u64 valid = 0;
    if (w0 >= block_base && w0 < hi) {
        u64 left = hi - w0;                 /* >= 1 */
        u64 nv = (left + 1) >> 1;           /* odd numbers in [w0, hi) */
        valid = nv >= 64 ? ~0ull : ((1ull << nv) - 1);
    }
    u64 bits = ~comp & valid;

    /* ---- wave-local compaction of candidate positions into LDS ---- */
    u32 cnt = __popcll(bits);
    u32 incl = cnt;
    for (int d = 1; d < 64; d <<= 1) {
        u32 v = __shfl_up(incl, d, 64);
        if (lane >= (u32)d) incl += v;
    }
    const u32 ncand = __shfl(incl, 63, 64);
    u32 pos = incl - cnt;
    uint16_t *wc = cand + wave * kWaveCandCap;
    while (bits) {
        u32 b = __ffsll((unsigned long long)bits) - 1;
        wc[pos++] = (uint16_t)(lane * 64 + b);
        bits &= bits - 1;
    }
    __syncthreads();

    /* ---- strong PRP over the compacted candidates, survivors compacted in place ---- */
    u32 nout = 0;
    for (u32 j0 = 0; j0 < ncand; j0 += 64) {
        u32 j = j0 + lane;
        bool act = j < ncand;
        u32 idx = act ? wc[j] : 0;
        bool prp = false;
        if (act) {
            u64 n = wave_base + 2ull * idx;
            u64 q = inv64(n);
            prp = spsp2(n, q, mont_one(n));
        }
        u64 bal = __ballot(prp);
        if (prp) {
            u32 below = __popcll(bal & ((1ull << lane) - 1));
            wc[nout + below] = (uint16_t)idx; /* nout + below <= j: already read */
        }
        nout += __popcll(bal);
    }

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

__device__ __forceinline__ void put_result(ww_result *res, u32 res_cap, DevState *st, u64 p,
                                           int32_t code, int32_t rem, int32_t quot)
{
    u32 ix = atomicAdd(&st->resultcount, 1u);
    if (ix < res_cap) {
        ww_result r;
        r.p = p;
        r.result = code;
        r.rem = rem;
        r.quot = quot;
        r.pad = 0;
        res[ix] = r;
    }
}

/* Result classification of wieferich.cl:247-293 (all comparisons unsigned 64-bit). */
__device__ __forceinline__ void classify_w(u64 p, u64 s0, u64 s1, ww_result *res, u32 cap,
                                           DevState *st)
{
    const u64 T = 1000; /* SPECIAL_THRESHOLD, WW.h:9 */
    if ((s1 == 0 && s0 == 1) || (s1 == p - 1 && s0 == p - 1)) {
        put_result(res, cap, st, p, 1093, 1, 0);
    } else if (s1 <= T) {
        if (s0 == 1) put_result(res, cap, st, p, 1093, 1, (int32_t)s1);
        else if (s0 == p - 1) put_result(res, cap, st, p, 1093, -1, (int32_t)(s1 + 1));
    } else if (s1 >= p - T) {
        if (s0 == 1) put_result(res, cap, st, p, 1093, 1, (int32_t)(u32)(s1 - p));
        else if (s0 == p - 1) put_result(res, cap, st, p, 1093, -1, (int32_t)(u32)(s1 + 1 - p));
    }
}

/* Result classification of wallsunsun.cl:325-349. */
__device__ __forceinline__ void classify_wss(u64 p, u64 c21, ww_result *res, u32 cap,
                                             DevState *st)
{
    const u64 T = 1000;
    if (c21 == 0) put_result(res, cap, st, p, 1, 0, 0);
    else if (c21 <= T) put_result(res, cap, st, p, 2, 0, (int32_t)c21);
    else if (c21 >= p - T) put_result(res, cap, st, p, 2, 0, (int32_t)(u32)(c21 - p));
}

__global__ __launch_bounds__(kTestThreads) void k_ww_tests(const u64 *list, u64 cap, DevState *st,
                                                            ww_result *res, u32 res_cap)
{
    const u32 count = st->primecount;
    const u64 n = count < cap ? count : cap;
    u64 sum = 0;
    for (u64 i = (u64)blockIdx.x * kTestThreads + threadIdx.x; i < n;
         i += (u64)gridDim.x * kTestThreads) {
        const u64 p = list[i];
        const u64 q = inv64(p);
        const m2 one = m2_one(p);
        const m2 s = wieferich_residue(p, q, one);
        const m2 f = wallsunsun_residue(p, q, one);
        sum += (s.d1 & 0xffffffffu) + (f.d1 & 0xffffffffu);
        classify_w(p, s.d0, s.d1, res, res_cap, st);
        classify_wss(p, f.d1, res, res_cap, st);
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

hipError_t launch_sieve_prp(u64 lo, u64 hi, u64 *list, u64 cap, DevState *st, hipStream_t s)
{
    u64 base = lo | 1;
    if (base >= hi) return hipSuccess;
    u64 span = hi - base;
    u64 blocks = (span + kIntsPerBlock - 1) / kIntsPerBlock;
    hipLaunchKernelGGL(k_sieve_prp, dim3((unsigned)blocks), dim3(kSieveThreads), 0, s, base, hi,
                       list, cap, st);
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
