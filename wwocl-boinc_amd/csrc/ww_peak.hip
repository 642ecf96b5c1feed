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

__global__ __launch_bounds__(256) void k_copy_probe(const uint4 *src, uint4 *dst, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
        dst[i] = src[i];
}

} // namespace

extern "C" int ww_copy_probe(int device, uint64_t bytes, double *ms)
{
    if (!ms || bytes < 16 || bytes % 16) return WW_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return WW_E_HIP;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return WW_E_HIP;
    uint4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess) return WW_E_HIP;
    if (hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return WW_E_HIP;
    }
    (void)hipMemset(a, 1, bytes);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(k_copy_probe, dim3(prop.multiProcessorCount * 8), dim3(256), 0, 0, a, b,
                       bytes / 16);
    (void)hipEventRecord(e1, 0);
    hipError_t e = hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return e == hipSuccess ? WW_OK : WW_E_HIP;
}

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
