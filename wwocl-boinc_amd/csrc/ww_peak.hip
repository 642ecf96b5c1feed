/*
 * ww_peak.hip -- issue-rate microbenchmark for the roofline of the search path.
 *
 * The path is bound by 32x32->64 integer multiplies on the VALU (v_mad_u64_u32); the
 * MI355X has no published rate for that instruction, so its peak is measured: eight
 * independent multiply-accumulate chains per lane on every CU. No counterpart in the
 * reference; used by bench.py to price the roofline.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ww_api.h"

namespace {

constexpr int kChains = 8;

__global__ __launch_bounds__(256) void k_mad_u64_u32(uint64_t *sink, uint32_t m, int iters)
{
    uint64_t acc[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) acc[j] = (uint64_t)(threadIdx.x + 1) * (j + 3) + blockIdx.x;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) acc[j] = (uint64_t)(uint32_t)acc[j] * m + acc[j];
    }
    uint64_t s = 0;
#pragma unroll
    for (int j = 0; j < kChains; ++j) s ^= acc[j];
    if (s == 0x9e3779b97f4a7c15ull) sink[0] = s; /* keeps the chains live */
}

__global__ __launch_bounds__(256) void k_mul_hi_u32(uint64_t *sink, uint32_t m, int iters)
{
    uint32_t acc[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) acc[j] = (threadIdx.x + 1) * (j + 3) + blockIdx.x;
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) acc[j] = __umulhi(acc[j], m) ^ acc[j];
    }
    uint32_t s = 0;
#pragma unroll
    for (int j = 0; j < kChains; ++j) s ^= acc[j];
    if (s == 0x9e3779b9u) sink[0] = s;
}

} // namespace

extern "C" int ww_peak_mul32(int device, int kind, double *ops_per_s)
{
    if (!ops_per_s || kind < 0 || kind > 1) return WW_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return WW_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return WW_E_HIP;
    uint64_t *sink = nullptr;
    if (hipMalloc(&sink, 8) != hipSuccess) return WW_E_HIP;
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    const int blocks = prop.multiProcessorCount * 8, threads = 256, iters = 8192;
    double best = 0;
    for (int rep = 0; rep < 4; ++rep) {
        (void)hipEventRecord(a, 0);
        if (kind == 0)
            hipLaunchKernelGGL(k_mad_u64_u32, dim3(blocks), dim3(threads), 0, 0, sink, 0x9e3779b9u, iters);
        else
            hipLaunchKernelGGL(k_mul_hi_u32, dim3(blocks), dim3(threads), 0, 0, sink, 0x9e3779b9u, iters);
        (void)hipEventRecord(b, 0);
        if (hipEventSynchronize(b) != hipSuccess) {
            (void)hipFree(sink);
            return WW_E_HIP;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, a, b);
        double ops = (double)blocks * threads * iters * kChains;
        if (rep > 0 && ms > 0 && ops / (ms * 1e-3) > best) best = ops / (ms * 1e-3);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(sink);
    *ops_per_s = best;
    return WW_OK;
}
