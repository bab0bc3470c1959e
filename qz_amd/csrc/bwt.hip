// Forward BWT on gfx950 by prefix-doubling suffix sorting, over a batch of
// independent blocks laid end to end in one position space.
//
// Replaces libbsc's bsc_bwt_encode (libsais / libcubwt; the reference links
// them from third_party, crates/qz-lib/build.rs:19-37,143) for the BSC block
// path of crates/qz-lib/src/compression/bsc.rs:153-191.
//
// Layout in HBM (positions are u32; a batch holds < 2^31 bytes):
//   text[N]      input bytes of all blocks, block b at [off[b], off[b+1])
//   sa[N]        suffix array, global positions, block b's suffixes at
//                SA slots [off[b], off[b+1])
//   rank[N]      for each text position, the SA slot where its current group
//                starts (Larsson-Sadakane ranks); the inverse SA at the end
//   act[m]       SA slots of suffixes still in groups of size >= 2
//   keys/vals    sort pairs for the active suffixes (ping-pong)
//
// Round 0 sorts all suffixes by (block id, first K0 symbols) with the
// symbols packed in mixed radix (sigma+1), 0 standing for "past the block
// end", so suffixes shorter than K0 are already distinguished. Round r >= 1
// sorts the active suffixes by (group start, rank[i+h] + 1), 0 when i+h hits
// the block end, h = K0 * 2^(r-1). Groups that become singletons leave the
// active list; the loop ends when it is empty.
#include "prims.h"
#include "engine.h"

#include <vector>
#include <algorithm>

namespace qz {

constexpr int kThreads = 256;

__device__ __forceinline__ uint32_t find_block(const uint32_t* __restrict__ off, uint32_t nblk, uint32_t i)
{
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (off[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(kThreads) void k_byte_hist(const uint8_t* __restrict__ text, uint32_t n,
                                                        uint32_t* __restrict__ hist)
{
    __shared__ uint32_t h[4][256];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE;
    for (int w = 0; w < 4; ++w) h[w][tid] = 0;
    __syncthreads();
    for (uint32_t i = blockIdx.x * kThreads + tid; i < n; i += gridDim.x * kThreads)
        atomicAdd(&h[wave][text[i]], 1u);
    __syncthreads();
    uint32_t t = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
    if (t) atomicAdd(&hist[tid], t);
}

struct PackParams {
    uint16_t code[256];  // symbol -> 1..sigma (sigma may be 256)
    uint64_t radix;      // sigma + 1
    uint32_t k0;         // symbols per key
    uint32_t cbits;      // key bits reserved for symbols; block id above
};

__global__ __launch_bounds__(kThreads) void k_round0_keys(const uint8_t* __restrict__ text, uint32_t n,
                                                          const uint32_t* __restrict__ off, uint32_t nblk,
                                                          PackParams pp, uint64_t* __restrict__ keys,
                                                          uint32_t* __restrict__ vals,
                                                          uint32_t* __restrict__ act)
{
    __shared__ uint16_t code[256];
    code[threadIdx.x] = pp.code[threadIdx.x];
    __syncthreads();
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= n) return;
    const uint32_t b = find_block(off, nblk, i);
    const uint32_t end = off[b + 1];
    uint64_t key = 0;
    for (uint32_t j = 0; j < pp.k0; ++j) {
        const uint32_t p = i + j;
        const uint32_t c = (p < end) ? code[text[p]] : 0u;
        key = key * pp.radix + c;
    }
    if (nblk > 1) key |= (uint64_t)b << pp.cbits;
    keys[i] = key;
    vals[i] = i;
    act[i] = i;
}

// hv[j] = act[j] at group heads, 0 elsewhere (for the max-scan).
__global__ __launch_bounds__(kThreads) void k_heads(const uint64_t* __restrict__ keys, uint32_t m,
                                                    const uint32_t* __restrict__ act, uint32_t* __restrict__ hv)
{
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    const bool head = (j == 0) || keys[j] != keys[j - 1];
    hv[j] = head ? act[j] : 0u;
}

// Writes back the sorted slice, updates ranks, flags suffixes still unsorted.
__global__ __launch_bounds__(kThreads) void k_apply(const uint64_t* __restrict__ keys,
                                                    const uint32_t* __restrict__ vals, uint32_t m,
                                                    const uint32_t* __restrict__ act,
                                                    const uint32_t* __restrict__ gstart,
                                                    uint32_t* __restrict__ sa, uint32_t* __restrict__ rank,
                                                    uint32_t* __restrict__ flag)
{
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    const uint64_t k = keys[j];
    const bool head = (j == 0) || k != keys[j - 1];
    const bool next_head = (j + 1 == m) || keys[j + 1] != k;
    const uint32_t s = vals[j];
    sa[act[j]] = s;
    rank[s] = gstart[j];
    flag[j] = (head && next_head) ? 0u : 1u;
}

// Compacts still-active SA slots; incl = inclusive scan of flags.
__global__ __launch_bounds__(kThreads) void k_compact(const uint32_t* __restrict__ incl,
                                                      const uint32_t* __restrict__ act_in, uint32_t m,
                                                      uint32_t* __restrict__ act_out)
{
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    const uint32_t p = incl[j];
    const uint32_t prev = j ? incl[j - 1] : 0u;
    if (p != prev) act_out[p - 1] = act_in[j];
}

// Keys of round h: (group start, rank[s+h] + 1 or 0 at the block end).
__global__ __launch_bounds__(kThreads) void k_round_keys(const uint32_t* __restrict__ act, uint32_t m,
                                                         const uint32_t* __restrict__ sa,
                                                         const uint32_t* __restrict__ rank,
                                                         const uint32_t* __restrict__ off, uint32_t nblk,
                                                         uint32_t h, uint64_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals)
{
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= m) return;
    const uint32_t s = sa[act[j]];
    const uint32_t end = off[find_block(off, nblk, s) + 1];
    const uint32_t g = rank[s];
    const uint64_t p = (uint64_t)s + h;
    const uint32_t k2 = (p < end) ? rank[(uint32_t)p] + 1u : 0u;
    keys[j] = ((uint64_t)g << 32) | k2;
    vals[j] = s;
}

// libsais-convention BWT of every block: U[0] = T[n-1]; U[j] = T[SA[j-1]-1]
// for 0 < j < p, T[SA[j]-1] for j >= p, with p = 1 + rank of suffix 0.
__global__ __launch_bounds__(kThreads) void k_bwt_out(const uint8_t* __restrict__ text,
                                                      const uint32_t* __restrict__ sa,
                                                      const uint32_t* __restrict__ rank,
                                                      const uint32_t* __restrict__ off, uint32_t nblk,
                                                      uint32_t n, uint8_t* __restrict__ out)
{
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    const uint32_t b = find_block(off, nblk, j);
    const uint32_t o = off[b], e = off[b + 1];
    const uint32_t local = j - o;
    const uint32_t p = rank[o] - o + 1u;
    uint8_t c;
    if (local == 0) c = text[e - 1];
    else c = text[(local < p ? sa[j - 1] : sa[j]) - 1u];
    out[j] = c;
}

// Per block: primary index and sampled indexes I[k] = 1 + rank(k*r).
__global__ void k_bwt_index(const uint32_t* __restrict__ rank, const uint32_t* __restrict__ off, uint32_t nblk,
                            int32_t* __restrict__ idx /* nblk x 256 */)
{
    const uint32_t b = blockIdx.x;
    if (b >= nblk) return;
    const uint32_t o = off[b], n = off[b + 1] - o;
    uint32_t r = n / 8;
    r |= r >> 1; r |= r >> 2; r |= r >> 4; r |= r >> 8; r |= r >> 16;
    r = (r >> 1) + 1;
    const uint32_t cnt = (n - 1) / r;
    for (uint32_t k = threadIdx.x; k <= cnt && k < 256; k += blockDim.x)
        idx[(size_t)b * 256 + k] = (int32_t)(rank[o + k * r] - o + 1u);
}

static int bitlen(uint64_t v)
{
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

static PackParams make_pack(const uint32_t* hist, uint32_t nblk)
{
    PackParams pp;
    int sigma = 0;
    for (int c = 0; c < 256; ++c) pp.code[c] = hist[c] ? (uint16_t)(++sigma) : 0;
    pp.radix = (uint64_t)sigma + 1;
    const int bb = nblk > 1 ? bitlen(nblk - 1) : 0;
    pp.cbits = 64 - bb;
    // largest k with radix^k <= 2^cbits
    unsigned __int128 lim = (unsigned __int128)1 << pp.cbits, v = 1;
    uint32_t k = 0;
    while (k < 64 && v * pp.radix <= lim) { v *= pp.radix; ++k; }
    pp.k0 = k ? k : 1;
    return pp;
}

int BwtEngine::reserve(uint64_t n, uint32_t nblk)
{
    if (n <= cap_ && nblk <= cap_blk_) return 0;
    release();
    const uint64_t c = std::max<uint64_t>(n, 1024);
    const uint32_t cb = std::max<uint32_t>(nblk, 1);
    size_t tmpw = std::max(radix_tmp_words(c), scan_tmp_words(c)) + 1024;
    QZ_HIP(hipMalloc(&text_, c + 64));
    QZ_HIP(hipMalloc(&out_, c + 64));
    QZ_HIP(hipMalloc(&sa_, c * 4));
    QZ_HIP(hipMalloc(&rank_, c * 4));
    QZ_HIP(hipMalloc(&act_, c * 4));
    QZ_HIP(hipMalloc(&bufA_, c * 4));
    QZ_HIP(hipMalloc(&bufB_, c * 4));
    QZ_HIP(hipMalloc(&keys0_, c * 8));
    QZ_HIP(hipMalloc(&keys1_, c * 8));
    QZ_HIP(hipMalloc(&vals0_, c * 4));
    QZ_HIP(hipMalloc(&vals1_, c * 4));
    QZ_HIP(hipMalloc(&tmp_, tmpw * 4));
    QZ_HIP(hipMalloc(&off_, (cb + 1) * 4));
    QZ_HIP(hipMalloc(&idx_, (size_t)cb * 256 * 4));
    QZ_HIP(hipMalloc(&hist_, 256 * 4));
    QZ_HIP(hipHostMalloc(&pinned_, 64));
    cap_ = c;
    cap_blk_ = cb;
    return 0;
}

void BwtEngine::release()
{
    void* ps[] = {text_, out_, sa_, rank_, act_, bufA_, bufB_, keys0_, keys1_, vals0_, vals1_, tmp_, off_, idx_, hist_};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (pinned_) (void)hipHostFree(pinned_);
    text_ = out_ = nullptr;
    sa_ = rank_ = act_ = bufA_ = bufB_ = vals0_ = vals1_ = tmp_ = off_ = hist_ = nullptr;
    keys0_ = keys1_ = nullptr;
    idx_ = nullptr;
    pinned_ = nullptr;
    cap_ = 0;
    cap_blk_ = 0;
}

// Sorts the suffixes of the batch in text (device); off_ holds the offsets.
int BwtEngine::sort_suffixes(const uint8_t* text, uint32_t n, uint32_t nblk, hipStream_t s)
{
    rounds_ = 0;
    // alphabet
    QZ_HIP(hipMemsetAsync(hist_, 0, 256 * 4, s));
    hipLaunchKernelGGL(k_byte_hist, dim3(std::min<uint32_t>(qz_div_up(n, kThreads), 2048)), dim3(kThreads), 0, s,
                       text, n, hist_);
    QZ_HIP(hipGetLastError());
    uint32_t hh[256];
    QZ_HIP(hipMemcpyAsync(hh, hist_, sizeof(hh), hipMemcpyDeviceToHost, s));
    QZ_HIP(hipStreamSynchronize(s));
    const PackParams pp = make_pack(hh, nblk);
    k0_ = pp.k0;

    const uint32_t g = qz_div_up(n, kThreads);
    hipLaunchKernelGGL(k_round0_keys, dim3(g), dim3(kThreads), 0, s, text, n, off_, nblk, pp, keys0_, vals0_, act_);
    QZ_HIP(hipGetLastError());

    // round-0 digit passes: symbol bits, then block id bits
    int shifts[16], np = 0;
    unsigned __int128 maxsym = 1;
    for (uint32_t k = 0; k < pp.k0; ++k) maxsym *= pp.radix;
    int symbits = 0;
    {
        unsigned __int128 v = maxsym - 1;
        while (v) { ++symbits; v >>= 1; }
    }
    for (int sh = 0; sh < symbits; sh += 8) shifts[np++] = sh;
    if (nblk > 1)
        for (int sh = (int)pp.cbits; sh < 64; sh += 8) shifts[np++] = sh;

    uint32_t m = n;
    uint32_t h = pp.k0;
    uint32_t* act = act_;
    uint32_t* fA = bufA_;
    uint32_t* fB = bufB_;
    for (int round = 0;; ++round) {
        if (round > 0) {
            hipLaunchKernelGGL(k_round_keys, dim3(qz_div_up(m, kThreads)), dim3(kThreads), 0, s, act, m, sa_, rank_,
                               off_, nblk, h, keys0_, vals0_);
            QZ_HIP(hipGetLastError());
            np = 0;
            const int b2 = bitlen(n), bg = bitlen(n - 1);
            for (int sh = 0; sh < b2; sh += 8) shifts[np++] = sh;
            for (int sh = 32; sh < 32 + bg; sh += 8) shifts[np++] = sh;
            h = (h > 0x40000000u) ? 0x80000000u : h * 2;
        }
        bool alt = false;
        QZ_TRY(radix_sort_pairs<uint64_t>(keys0_, vals0_, keys1_, vals1_, m, shifts, np, tmp_, s, &alt));
        uint64_t* keys = alt ? keys1_ : keys0_;
        uint32_t* vals = alt ? vals1_ : vals0_;
        const uint32_t gm = qz_div_up(m, kThreads);
        hipLaunchKernelGGL(k_heads, dim3(gm), dim3(kThreads), 0, s, keys, m, act, fA);
        QZ_HIP(hipGetLastError());
        QZ_TRY((device_scan<OpMax, true>(fA, fA, m, tmp_, s)));
        hipLaunchKernelGGL(k_apply, dim3(gm), dim3(kThreads), 0, s, keys, vals, m, act, fA, sa_, rank_, fB);
        QZ_HIP(hipGetLastError());
        QZ_TRY((device_scan<OpAdd, true>(fB, fB, m, tmp_, s)));
        hipLaunchKernelGGL(k_compact, dim3(gm), dim3(kThreads), 0, s, fB, act, m, fA);
        QZ_HIP(hipGetLastError());
        QZ_HIP(hipMemcpyAsync(pinned_, fB + (m - 1), 4, hipMemcpyDeviceToHost, s));
        QZ_HIP(hipStreamSynchronize(s));
        const uint32_t m2 = *(volatile uint32_t*)pinned_;
        ++rounds_;
        // rotate: fA holds the new active list
        uint32_t* t = act;
        act = fA;
        fA = t;
        m = m2;
        if (m == 0) break;
        if (rounds_ > 40) {
            fprintf(stderr, "qzbsc: suffix sort did not converge\n");
            return QZ_GPU_ERROR;
        }
    }
    return 0;
}

int BwtEngine::encode_device(const uint8_t* text, uint8_t* out, uint32_t n, uint32_t nblk, hipStream_t s)
{
    QZ_TRY(sort_suffixes(text, n, nblk, s));
    hipLaunchKernelGGL(k_bwt_out, dim3(qz_div_up(n, kThreads)), dim3(kThreads), 0, s, text, sa_, rank_, off_, nblk, n,
                       out);
    QZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_bwt_index, dim3(nblk), dim3(64), 0, s, rank_, off_, nblk, idx_);
    QZ_HIP(hipGetLastError());
    return 0;
}

}  // namespace qz
