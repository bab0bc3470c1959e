// Device-wide primitives used by the block-sorting kernels: multi-level scans
// and an LSD radix sort with wave-ballot stable ranking. Everything here is
// written for 64-lane wavefronts and 256-thread workgroups (4 waves).
#pragma once
#include "common.h"

namespace qz {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanThreads * kScanItems;  // 1024

struct OpAdd {
    static __device__ __forceinline__ uint32_t apply(uint32_t a, uint32_t b) { return a + b; }
    static constexpr uint32_t identity = 0u;
};
struct OpMax {
    static __device__ __forceinline__ uint32_t apply(uint32_t a, uint32_t b) { return a > b ? a : b; }
    static constexpr uint32_t identity = 0u;
};

// Inclusive scan of one value per lane across the wave.
template <class Op>
__device__ __forceinline__ uint32_t wave_inclusive(uint32_t v)
{
    const uint32_t lane = lane_id();
#pragma unroll
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        uint32_t o = __shfl_up(v, d, QZ_WAVE);
        if (lane >= (uint32_t)d) v = Op::apply(o, v);
    }
    return v;
}

// Block-level scan of kScanTile items; writes local scan and the block total.
template <class Op, bool kInclusive>
__global__ __launch_bounds__(kScanThreads) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                             uint32_t* __restrict__ out, uint32_t n,
                                                             uint32_t* __restrict__ sums)
{
    __shared__ uint32_t wsum[kScanThreads / QZ_WAVE];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, lane = tid % QZ_WAVE;
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)tid * kScanItems;
    uint32_t v[kScanItems];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) v[i] = (base + i < n) ? in[base + i] : Op::identity;
    uint32_t acc = Op::identity;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) acc = Op::apply(acc, v[i]);
    uint32_t incl = wave_inclusive<Op>(acc);
    if (lane == QZ_WAVE - 1) wsum[wave] = incl;
    __syncthreads();
    uint32_t wpre = Op::identity;
    for (uint32_t w = 0; w < wave; ++w) wpre = Op::apply(wpre, wsum[w]);
    uint32_t excl_thread = __shfl_up(incl, 1, QZ_WAVE);
    if (lane == 0) excl_thread = Op::identity;
    uint32_t run = Op::apply(wpre, excl_thread);
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        uint32_t nr = Op::apply(run, v[i]);
        if (base + i < n) out[base + i] = kInclusive ? nr : run;
        run = nr;
    }
    if (sums && tid == kScanThreads - 1) {
        uint32_t tot = Op::identity;
        for (uint32_t w = 0; w < kScanThreads / QZ_WAVE; ++w) tot = Op::apply(tot, wsum[w]);
        sums[blockIdx.x] = tot;
    }
}

template <class Op>
__global__ __launch_bounds__(kScanThreads) void k_scan_apply(uint32_t* __restrict__ out, uint32_t n,
                                                             const uint32_t* __restrict__ pre)
{
    const uint32_t p = pre[blockIdx.x];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    for (uint32_t i = threadIdx.x; i < kScanTile; i += kScanThreads)
        if (base + i < n) out[base + i] = Op::apply(p, out[base + i]);
}

// Scratch words a scan of n elements needs.
static inline size_t scan_tmp_words(uint64_t n)
{
    size_t words = 0;
    while (n > (uint64_t)kScanTile) {
        n = (n + kScanTile - 1) / kScanTile;
        words += n + 1;
    }
    return words + 1;
}

// out may alias in. If total != nullptr the grand total (sum or max) is
// written to device memory *total.
template <class Op, bool kInclusive>
static int device_scan(const uint32_t* in, uint32_t* out, uint64_t n, uint32_t* tmp, hipStream_t s)
{
    if (n == 0) return 0;
    const uint32_t tiles = qz_div_up(n, kScanTile);
    if (tiles == 1) {
        hipLaunchKernelGGL((k_scan_tiles<Op, kInclusive>), dim3(1), dim3(kScanThreads), 0, s, in, out,
                           (uint32_t)n, nullptr);
        QZ_HIP(hipGetLastError());
        return 0;
    }
    uint32_t* sums = tmp;
    hipLaunchKernelGGL((k_scan_tiles<Op, kInclusive>), dim3(tiles), dim3(kScanThreads), 0, s, in, out,
                       (uint32_t)n, sums);
    QZ_HIP(hipGetLastError());
    // exclusive scan of tile totals, then apply
    QZ_TRY((device_scan<Op, false>(sums, sums, tiles, tmp + tiles + 1, s)));
    hipLaunchKernelGGL((k_scan_apply<Op>), dim3(tiles), dim3(kScanThreads), 0, s, out, (uint32_t)n, sums);
    QZ_HIP(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Block-segmented LSD radix sort of (key, u32 value) pairs with BITS-bit
// digits (8 or 11). Tiles of kSortTile never span two blocks of a batch;
// histograms are laid out [block][digit][tile], so one exclusive scan over
// them yields, for every (tile, digit), the global output slot: block b's
// elements stay inside [off[b], off[b+1]) and no block id bits are needed
// in the keys. Stable.
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 256;
constexpr int kSortTile = 4096;

// Block-aligned tiles; returns the histogram words (radix per tile).
static inline uint32_t build_tiles(const uint32_t* offs, uint32_t nblk, TileDesc* out, uint32_t* count,
                                   uint32_t radix = 256)
{
    uint32_t hbase = 0, k = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t o = offs[b], nb = offs[b + 1] - o;
        const uint32_t nt = qz_div_up(nb, kSortTile);
        for (uint32_t t = 0; t < nt; ++t) {
            TileDesc d;
            d.start = o + t * kSortTile;
            d.len = (nb - t * kSortTile) < (uint32_t)kSortTile ? nb - t * kSortTile : (uint32_t)kSortTile;
            d.block = b;
            d.hbase = hbase;
            d.ntiles = nt;
            d.tlocal = t;
            out[k++] = d;
        }
        hbase += nt * radix;
    }
    *count = k;
    return hbase;
}

// The tile count and histogram words build_tiles would produce, in O(nblk).
static inline uint32_t count_tiles(const uint32_t* offs, uint32_t nblk, uint32_t* count, uint32_t radix = 256)
{
    uint32_t k = 0;
    for (uint32_t b = 0; b < nblk; ++b) k += qz_div_up(offs[b + 1] - offs[b], kSortTile);
    *count = k;
    return k * radix;
}

static inline uint32_t max_tiles(uint64_t n, uint32_t nblk) { return (uint32_t)(n / kSortTile + nblk + 1); }

template <typename K, int BITS>
__global__ __launch_bounds__(kSortThreads) void k_seg_hist(const K* __restrict__ keys,
                                                           const TileDesc* __restrict__ tiles, int shift,
                                                           uint32_t* __restrict__ hist)
{
    constexpr uint32_t D = 1u << BITS;
    constexpr int W = kSortThreads / QZ_WAVE;
    constexpr int IT = kSortTile / kSortThreads;
    // R copies of each wave's counters, a digit's copies in adjacent words
    // (different banks), the copy chosen by lane: lanes of one wave that
    // share a digit (frequent with 2-bit symbols packed into digits) then
    // mostly update different words instead of queueing on one.
    constexpr uint32_t R = BITS <= 9 ? 4u : 1u;
    __shared__ uint32_t cnt[W][D * R];
    const TileDesc td = tiles[xcd_item(blockIdx.x, gridDim.x)];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, rep = tid & (R - 1);
    if (td.len == 0) {  // uniform; build_tiles makes no empty tiles
        for (uint32_t d = tid; d < D; d += kSortThreads) hist[(size_t)td.hbase + (size_t)d * td.ntiles + td.tlocal] = 0u;
        return;
    }
    for (uint32_t d = tid; d < D * R; d += kSortThreads)
#pragma unroll
        for (int w = 0; w < W; ++w) cnt[w][d] = 0;
    // The counters are cleared before any load is issued (a workgroup
    // barrier waits for every load in flight); then all of the tile's loads
    // go out at once and are counted as they arrive. Lanes past the tile's
    // end read its last key and count nothing.
    __syncthreads();
    const uint32_t last = td.start + td.len - 1u;
    K kk[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) kk[it] = keys[min(td.start + tid + (uint32_t)it * kSortThreads, last)];
#pragma unroll
    for (int it = 0; it < IT; ++it)
        if (tid + (uint32_t)it * kSortThreads < td.len)
            atomicAdd(&cnt[wave][((uint32_t)(kk[it] >> shift) & (D - 1)) * R + rep], 1u);
    __syncthreads();
    for (uint32_t d = tid; d < D; d += kSortThreads) {
        uint32_t t = 0;
#pragma unroll
        for (int w = 0; w < W; ++w)
#pragma unroll
            for (uint32_t r = 0; r < R; ++r) t += cnt[w][d * R + r];
        hist[(size_t)td.hbase + (size_t)d * td.ntiles + td.tlocal] = t;
    }
}

// One LSD pass over a tile of <= 4096 pairs. Each wave ranks its own 1024
// consecutive elements (16 iterations of 64) with wave-private LDS counters;
// the tile is then permuted by digit in LDS and written out so that
// consecutive lanes store consecutive slots of a digit run. Stable: element
// order is wave-major, then iteration, then lane.
// Ranking an iteration needs, per lane, the mask of lanes holding the same
// digit. For digits of <= 9 bits each lane ORs its own bit into a per-wave
// LDS word of its digit and reads the word back (one 64-bit LDS atomic and
// one load, in place of one ballot and four 32-bit mask operations per
// digit bit); the digit's first lane clears the word for the next
// iteration. A wave's LDS instructions complete in order, so every lane
// reads the word after all of the iteration's ORs. Wider digits use one
// ballot per bit.
// VS: where a pass takes its values from. VFromArray reads them; a source
// may instead compute them from the element's position (the first pass of a
// sort whose values are a function of position). A source provides its LDS
// (Lds), fills it before the pass's first barrier (stage) and yields the
// values of a lane's IT elements (load; lanes past the tile's end get the
// last element's value).
struct VFromArray {
    const uint32_t* v;
    struct Lds {};
    __device__ __forceinline__ void stage(Lds&) const {}
    template <int IT>
    __device__ __forceinline__ void load(const Lds&, const TileDesc& td, uint32_t e0, uint32_t last,
                                         uint32_t (&vv)[IT]) const
    {
#pragma unroll
        for (int it = 0; it < IT; ++it) vv[it] = v[min(td.start + e0 + (uint32_t)it * QZ_WAVE, last)];
    }
};

template <typename K, int BITS, class VS>
__global__ __launch_bounds__(kSortThreads) void k_seg_scatter(const K* __restrict__ kin, VS vs,
                                                              K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              const TileDesc* __restrict__ tiles, int shift,
                                                              const uint32_t* __restrict__ offs,
                                                              const int32_t* __restrict__ seg_delta)
{
    constexpr uint32_t D = 1u << BITS;
    constexpr uint32_t R = D / kSortThreads;  // digits per thread in the prefix (1 or 8)
    constexpr int W = kSortThreads / QZ_WAVE;
    constexpr int IT = kSortTile / kSortThreads;  // 16
    constexpr bool kOrMask = (size_t)W * D * sizeof(uint64_t) <= sizeof(K) * kSortTile;
    __shared__ __align__(16) K sk[kSortTile];
    __shared__ uint32_t sv[kSortTile];
    __shared__ uint32_t wcnt[W][D];
    __shared__ uint32_t lstart[D];
    __shared__ uint32_t gbase[D];
    __shared__ uint32_t wtot[W];
    __shared__ typename VS::Lds vl;
    // [W][D] mask words. For 8-bit digits in an array of their own, which
    // also keeps the workgroup's LDS above 40 KB, so at most 3 workgroups
    // share a CU: with 4 a pass over 1.5e9 pairs took 8.4 ms, with 3 6.0 ms.
    // With 9-bit digits the other arrays already hold 44 KB, and the masks
    // share the key staging array, which is idle until the permutation.
    constexpr bool kOwnMask = kOrMask && BITS <= 8;
    __shared__ unsigned long long mown[kOwnMask ? W * D : 1];
    uint64_t* const msk = kOwnMask ? reinterpret_cast<uint64_t*>(mown) : reinterpret_cast<uint64_t*>(sk);
    const TileDesc td = tiles[xcd_item(blockIdx.x, gridDim.x)];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, lane = tid % QZ_WAVE;
    if (td.len == 0) return;  // uniform; build_tiles makes no empty tiles
    for (uint32_t d = tid; d < D; d += kSortThreads) {
#pragma unroll
        for (int w = 0; w < W; ++w) wcnt[w][d] = 0;
        if (kOrMask)
#pragma unroll
            for (int w = 0; w < W; ++w) msk[(uint32_t)w * D + d] = 0ull;
    }
    vs.stage(vl);
    // The counters are cleared before any load is issued: a workgroup
    // barrier waits for every load in flight, so the tile's loads go out
    // after it and the ranking consumes them as they arrive. Lanes past the
    // tile's end read its last element and rank nothing.
    __syncthreads();
    const uint32_t e0 = wave * (IT * QZ_WAVE) + lane;
    const uint32_t last = td.start + td.len - 1u;
    K kk[IT];
    uint32_t vv[IT];
    uint32_t rk[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) kk[it] = kin[min(td.start + e0 + (uint32_t)it * QZ_WAVE, last)];
    vs.load(vl, td, e0, last, vv);
    // this tile's output start per digit, stored to LDS after the ranking;
    // seg_delta (optional): per segment, where its output lies relative to
    // the packed order the histograms count in
    const uint32_t delta = seg_delta ? (uint32_t)seg_delta[td.block] : 0u;
    uint32_t gb[R];
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) gb[r] = offs[(size_t)td.hbase + (size_t)(tid * R + r) * td.ntiles + td.tlocal] + delta;
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const bool valid = e0 + (uint32_t)it * QZ_WAVE < td.len;
        const uint32_t d = (uint32_t)(kk[it] >> shift) & (D - 1);
        uint64_t peer;
        if (kOrMask) {
            uint64_t* const word = msk + wave * D + d;
            if (valid) atomicOr((unsigned long long*)word, 1ull << lane);
            peer = valid ? *word : 0ull;
        } else {
            peer = __ballot(valid);
#pragma unroll
            for (int b = 0; b < BITS; ++b) {
                const bool bit = (d >> b) & 1u;
                const uint64_t m = __ballot(bit);
                peer &= bit ? m : ~m;
            }
        }
        const uint32_t below = popc_below(peer);
        const uint32_t old = wcnt[wave][d];
        if (valid && below == 0) {
            wcnt[wave][d] = old + (uint32_t)__popcll(peer);
            if (kOrMask) msk[wave * D + d] = 0ull;
        }
        rk[it] = valid ? (old + below) : 0xffffffffu;
    }
#pragma unroll
    for (uint32_t r = 0; r < R; ++r) gbase[tid * R + r] = gb[r];
    __syncthreads();
    {
        // digits [tid*R, tid*R+R): wave offsets per digit, then the tile-local
        // start of every digit run by a block scan of per-thread totals
        uint32_t tot = 0;
        uint32_t dtot[R];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            const uint32_t d = tid * R + r;
            uint32_t acc = 0;
#pragma unroll
            for (int w = 0; w < W; ++w) {
                const uint32_t c = wcnt[w][d];
                wcnt[w][d] = acc;
                acc += c;
            }
            dtot[r] = acc;
            tot += acc;
        }
        const uint32_t incl = wave_inclusive<OpAdd>(tot);
        if (lane == QZ_WAVE - 1) wtot[wave] = incl;
        __syncthreads();
        uint32_t run = incl - tot;
        for (uint32_t w = 0; w < wave; ++w) run += wtot[w];
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) {
            lstart[tid * R + r] = run;
            run += dtot[r];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (rk[it] != 0xffffffffu) {
            const uint32_t d = (uint32_t)(kk[it] >> shift) & (D - 1);
            const uint32_t pos = lstart[d] + wcnt[wave][d] + rk[it];
            sk[pos] = kk[it];
            sv[pos] = vv[it];
        }
    }
    __syncthreads();
    for (uint32_t i = tid; i < td.len; i += kSortThreads) {
        const K k = sk[i];
        const uint32_t d = (uint32_t)(k >> shift) & (D - 1);
        const uint32_t o = gbase[d] + (i - lstart[d]);
        kout[o] = k;
        vout[o] = sv[i];
    }
}

// Sorts by key bits [0, key_bits) with BITS-bit digits. `tiles` must have
// been built with radix 1 << BITS (hist_words from build_tiles). hist holds
// hist_words, stmp the scan scratch (scan_tmp_words(hist_words)). Result in (k0,v0) for even pass counts.
// first: the first pass's values (VFromArray{v0} to read them from v0; v0
// then serves only as a buffer of the later passes).
// kin (optional): the first pass reads its keys from kin instead of k0,
// which then serves only as a buffer, so kin is left as it was.
// seg_delta (optional): segment b's elements are at tile positions (tiles'
// start) and are sorted into positions shifted by seg_delta[b] from the
// packed order of the histograms (segments laid out with gaps).
template <typename K, int BITS, class VS0>
static int seg_radix_sort(K* k0, uint32_t* v0, K* k1, uint32_t* v1, const TileDesc* tiles, uint32_t ntiles,
                          uint32_t hist_words, int key_bits, uint32_t* hist, uint32_t* stmp, hipStream_t s,
                          bool* in_alt, VS0 first, const K* kin = nullptr, const int32_t* seg_delta = nullptr)
{
    *in_alt = false;
    K *ka = k0, *kb = k1;
    uint32_t *va = v0, *vb = v1;
    int passes = 0;
    for (int sh = 0; sh < key_bits; sh += BITS, ++passes) {
        const K* src = (passes == 0 && kin) ? kin : ka;
        hipLaunchKernelGGL((k_seg_hist<K, BITS>), dim3(ntiles), dim3(kSortThreads), 0, s, src, tiles, sh, hist);
        QZ_HIP(hipGetLastError());
        QZ_TRY((device_scan<OpAdd, false>(hist, hist, hist_words, stmp, s)));
        if (passes == 0)
            hipLaunchKernelGGL((k_seg_scatter<K, BITS, VS0>), dim3(ntiles), dim3(kSortThreads), 0, s, src, first, kb,
                               vb, tiles, sh, hist, seg_delta);
        else
            hipLaunchKernelGGL((k_seg_scatter<K, BITS, VFromArray>), dim3(ntiles), dim3(kSortThreads), 0, s, src,
                               VFromArray{va}, kb, vb, tiles, sh, hist, seg_delta);
        QZ_HIP(hipGetLastError());
        K* t = ka; ka = kb; kb = t;
        uint32_t* tv = va; va = vb; vb = tv;
    }
    *in_alt = (passes & 1) != 0;
    return 0;
}

// Largest digit width the sorts use (sizes histogram scratch).
constexpr int kMaxRadixBits = 11;

}  // namespace qz
