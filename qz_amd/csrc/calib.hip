// Memory-system calibration kernels (not on the block path).
//
// The suffix sort and the inverse BWT are bound by scattered 4-byte loads
// and stores, not by streaming bandwidth. These kernels measure what the
// device sustains for exactly those access shapes over an array of a given
// size, so each random-access kernel of the engine can be read against the
// same shape's ceiling:
//   gather   dst[i] = src[idx[i]]   idx and dst streamed, one scattered load
//                                   (the shape of the doubling rounds' key
//                                   gathers and of the BWT output gathers)
//   scatter  dst[idx[i]] = i        idx streamed, one scattered store (the
//                                   shape of the rank updates)
//   copy     dst[i] = src[i]        16-byte vectors, the streaming reference
// Indexes are uniform over the array (a multiplicative hash of i).
#include "common.h"
#include "../../include/qzbsc.h"

#include <algorithm>

namespace qz {

constexpr int kCalThreads = 256;

__device__ __forceinline__ uint32_t cal_mix(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

__global__ __launch_bounds__(kCalThreads) void k_cal_index(uint32_t* __restrict__ idx, uint64_t items, uint64_t m,
                                                           uint32_t seed)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCalThreads + threadIdx.x; i < items;
         i += (uint64_t)gridDim.x * kCalThreads)
        idx[i] = (uint32_t)(((uint64_t)cal_mix((uint32_t)i ^ seed) * m) >> 32);
}

__global__ __launch_bounds__(kCalThreads) void k_cal_iota(uint32_t* __restrict__ a, uint64_t m)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCalThreads + threadIdx.x; i < m; i += (uint64_t)gridDim.x * kCalThreads)
        a[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kCalThreads) void k_cal_gather(const uint32_t* __restrict__ src,
                                                            const uint32_t* __restrict__ idx,
                                                            uint32_t* __restrict__ dst, uint64_t items)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCalThreads + threadIdx.x; i < items;
         i += (uint64_t)gridDim.x * kCalThreads)
        dst[i] = src[idx[i]];
}

__global__ __launch_bounds__(kCalThreads) void k_cal_scatter(uint32_t* __restrict__ dst,
                                                             const uint32_t* __restrict__ idx, uint64_t items)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCalThreads + threadIdx.x; i < items;
         i += (uint64_t)gridDim.x * kCalThreads)
        dst[idx[i]] = (uint32_t)i;
}

__global__ __launch_bounds__(kCalThreads) void k_cal_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                          uint64_t n16)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kCalThreads + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * kCalThreads)
        dst[i] = src[i];
}

static int calib_run(int mode, uint64_t array_bytes, uint64_t items, int iters, double* ms)
{
    const uint64_t m = array_bytes / 4;
    if (m < 1024 || m > 0xffffffffull || iters < 1 || (mode != 2 && (items < 1 || items > 0xffffffffull)))
        return QZ_BAD_PARAMETER;
    uint32_t *src = nullptr, *dst = nullptr, *idx = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    auto run = [&]() -> int {
        QZ_HIP(hipStreamCreate(&s));
        QZ_HIP(hipEventCreate(&e0));
        QZ_HIP(hipEventCreate(&e1));
        QZ_HIP(hipMalloc(&src, m * 4));
        // gather: dst holds `items` words; scatter: dst is the array; copy: a second array
        QZ_HIP(hipMalloc(&dst, (mode == 0 ? items : m) * 4));
        if (mode != 2) QZ_HIP(hipMalloc(&idx, items * 4));
        const uint32_t grid = 256u * 16u;  // 16 workgroups per CU, grid-stride
        hipLaunchKernelGGL(k_cal_iota, dim3(grid), dim3(kCalThreads), 0, s, src, m);
        if (mode != 2) hipLaunchKernelGGL(k_cal_index, dim3(grid), dim3(kCalThreads), 0, s, idx, items, m, 0x9e3779b9u);
        QZ_HIP(hipGetLastError());
        auto once = [&]() {
            if (mode == 0)
                hipLaunchKernelGGL(k_cal_gather, dim3(grid), dim3(kCalThreads), 0, s, src, idx, dst, items);
            else if (mode == 1)
                hipLaunchKernelGGL(k_cal_scatter, dim3(grid), dim3(kCalThreads), 0, s, dst, idx, items);
            else
                hipLaunchKernelGGL(k_cal_copy, dim3(grid), dim3(kCalThreads), 0, s, (const uint4*)src, (uint4*)dst,
                                   m / 4);
        };
        once();  // warm the TLB and caches once
        QZ_HIP(hipGetLastError());
        QZ_HIP(hipEventRecord(e0, s));
        for (int k = 0; k < iters; ++k) once();
        QZ_HIP(hipEventRecord(e1, s));
        QZ_HIP(hipGetLastError());
        QZ_HIP(hipEventSynchronize(e1));
        float t = 0;
        QZ_HIP(hipEventElapsedTime(&t, e0, e1));
        *ms = (double)t / iters;
        return 0;
    };
    rc = run();
    if (s) (void)hipStreamSynchronize(s);
    for (void* p : {(void*)src, (void*)dst, (void*)idx})
        if (p) (void)hipFree(p);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    return rc;
}

}  // namespace qz

extern "C" QZBSC_API int qzbsc_calib_access(int mode, long long array_bytes, long long items, int iters, double* ms)
{
    if (!ms || mode < 0 || mode > 2 || array_bytes <= 0 || items < 0) return QZ_BAD_PARAMETER;
    return qz::calib_run(mode, (uint64_t)array_bytes, (uint64_t)items, iters, ms);
}
