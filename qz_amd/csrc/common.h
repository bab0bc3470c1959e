// Shared helpers for the gfx950 block-sorting kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#define QZ_WAVE 64

// Error codes follow libbsc's (crates/qz-lib/src/compression/bsc.rs:18-24);
// GPU runtime failures map to -7 (below libbsc's range).
#define QZ_NO_ERROR 0
#define QZ_BAD_PARAMETER -1
#define QZ_NOT_ENOUGH_MEMORY -2
#define QZ_NOT_COMPRESSIBLE -3
#define QZ_NOT_SUPPORTED -4
#define QZ_UNEXPECTED_EOB -5
#define QZ_DATA_CORRUPT -6
#define QZ_GPU_ERROR -7

#define QZ_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t _e = (call);                                                               \
        if (_e != hipSuccess) {                                                               \
            fprintf(stderr, "qzbsc: HIP error %s at %s:%d\n", hipGetErrorString(_e), __FILE__, \
                    __LINE__);                                                                \
            return QZ_GPU_ERROR;                                                              \
        }                                                                                     \
    } while (0)

#define QZ_TRY(expr)              \
    do {                          \
        int _rc = (expr);         \
        if (_rc != 0) return _rc; \
    } while (0)

namespace qz {
// A radix/ranking tile that never spans two blocks of a batch.
struct TileDesc {
    uint32_t start;   // first element (global)
    uint32_t len;     // elements in the tile
    uint32_t block;
    uint32_t hbase;   // base of this block's [digit][tile] histogram region
    uint32_t ntiles;  // tiles of this block
    uint32_t tlocal;  // tile index within the block
};
}  // namespace qz

static inline uint32_t qz_div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

__device__ __forceinline__ uint32_t lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u)); }

// Number of set bits of `mask` in lanes strictly below this lane.
__device__ __forceinline__ uint32_t popc_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

namespace qz {
// Workgroups go to the 8 XCDs of the MI355X round-robin by workgroup id, and
// each XCD has its own L2. xcd_item maps workgroup `bid` of a grid of n to an
// item index so that every XCD takes one contiguous range of items:
// neighbouring tiles, which share histogram and output lines, then meet in
// one L2 instead of eight. A bijection on [0, n).
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_item(uint32_t bid, uint32_t n)
{
    const uint32_t x = bid % kXcds, j = bid / kXcds, q = n / kXcds, r = n % kXcds;
    return x * q + (x < r ? x : r) + j;
}

// QZBSC_TRACE_MEM=1 prints every workspace reallocation and its duration.
struct MemTrace {
    const char* what;
    unsigned long long n;
    unsigned nblk;
    std::chrono::steady_clock::time_point t0;
    bool on;
    MemTrace(const char* w, unsigned long long n_, unsigned nb)
        : what(w), n(n_), nblk(nb), t0(std::chrono::steady_clock::now()), on(getenv("QZBSC_TRACE_MEM") != nullptr)
    {
    }
    ~MemTrace()
    {
        if (!on) return;
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        fprintf(stderr, "qzbsc: %s workspace regrown for %llu bytes, %u blocks in %.1f ms\n", what, n, nblk, ms);
    }
};
}  // namespace qz
