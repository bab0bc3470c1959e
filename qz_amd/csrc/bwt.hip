// Forward BWT on gfx950 by prefix-doubling suffix sorting, over a batch of
// independent blocks laid end to end in one position space.
//
// Replaces libbsc's bsc_bwt_encode (libsais / libcubwt; the reference links
// them from third_party, crates/qz-lib/build.rs:19-37,143) for the BSC block
// path of crates/qz-lib/src/compression/bsc.rs:153-191.
//
// Layout in HBM (positions are u32; a batch holds < 2^31 bytes):
//   text[N]      input bytes of all blocks, block b at [off[b], off[b+1])
//   sa[N]        suffix array; block b's suffixes occupy SA slots
//                [off[b], off[b+1]); an entry is the suffix's global
//                position, or (char mode, see SaVal) its block-local position
//                and the code of the symbol before it
//   rank[N]      per text position, the SA slot where its current group
//                starts (Larsson-Sadakane ranks); the inverse SA at the end
//   key2[N]      per SA slot, the doubling key gathered for this round
//   groups       unsorted groups (start slot, size) in eight size classes:
//                2, 3-4, 5-8, 9-16, 17-32, 33-64, 65-4096, > 4096
//
// Round 0 radix-sorts all suffixes by (block id, first K0 symbols), the
// symbols packed in mixed radix (sigma+1) with 0 for "past the block end",
// so suffixes shorter than K0 are already distinguished. Every later round
// (h = K0 * 2^(r-1)) first gathers key2 = rank[s+h] + 1 (0 at the block
// end) for every unsorted suffix, then sorts each group by key2:
//   * groups of <= 64 inside one wavefront (bitonic network in registers,
//     64 / 2^c groups per wave),
//   * groups of 65..4096 inside one workgroup (bitonic network in LDS),
//   * larger groups with the device radix sort on (group id, key2).
// Sorting a group splits it into subgroups; subgroups of size >= 2 go to
// the next round's lists. The loop ends when every list is empty.
#include "prims.h"
#include "engine.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

namespace qz {

constexpr int kThreads = 256;
constexpr int kClasses = 9;
constexpr int kSmallClasses = 6;  // sizes 2 .. 64
constexpr int kClsWave = 6;       // 65 .. 1024: one wavefront per group
constexpr int kClsMed = 7;        // 1025 .. 4096: one workgroup per group
constexpr int kClsLarge = 8;      // > 4096: segmented radix sort
constexpr int kWaveMax = 1024;
constexpr int kMedMax = 4096;
constexpr uint32_t kPadKey = 0xffffffffu;
// Device counters (group-list class counts, pool tops, bucket fills) each
// own a 256-byte line: atomics on words of one line serialise, and every
// workgroup that emits groups adds to several of them.
constexpr int kCtrStride = 64;

struct Grp {
    uint32_t start, size;
};

// Where a round emits its new unsorted groups: one pool in any order (at
// most n/2 groups of size >= 2 exist, so n/2 entries always suffice), with
// per-class counters. Before the next round k_bucket moves the pool into one
// list ordered by class, each class a contiguous run, which the next round's
// kernels read. The two counter sets (rounds alternate between them) fill
// the pool from opposite ends: set A upwards from entry 0, set B downwards
// from entry cap - 1. Round 0 can then emit into both at once (the second
// set takes groups that already went through the first doubling round), and
// as all live groups are disjoint and have >= 2 members, the two never meet.
struct GroupLists {
    Grp* pool;
    uint32_t* top;        // pool fill counter
    uint32_t* cnt;        // kClasses per-class counters
    uint32_t cap;         // pool entries
    uint32_t* overflow;   // set if the pool would overflow (cannot happen by construction)
    uint32_t down;        // 1: entry k of this set lives at pool[cap - 1 - k]
};

// Entry `at` of a set (counted in emission order).
__device__ __forceinline__ void pool_put(const GroupLists& gl, uint32_t at, uint32_t start, uint32_t size)
{
    if (at < gl.cap) gl.pool[gl.down ? gl.cap - 1u - at : at] = Grp{start, size};
    else *gl.overflow = 1u;
}

// Occupancy caps of the random-access kernels, as extra (unused) LDS per
// workgroup: beyond a few workgroups per CU their scattered 4-byte loads
// and stores only thrash the caches (swept on the MI355X; QZBSC_LDS_<name>
// overrides one, in bytes).
enum OccKernel { kOccSortSmall, kOccGatherSmall, kOccFinal0, kOccBwtOut, kOccFinalLarge, kOccCount };
constexpr uint32_t kOccDefault[kOccCount] = {16384, 0, 0, 0, 0};
static const char* const kOccEnv[kOccCount] = {"QZBSC_LDS_SORT_SMALL", "QZBSC_LDS_GATHER_SMALL", "QZBSC_LDS_FINAL0",
                                               "QZBSC_LDS_BWT_OUT", "QZBSC_LDS_FINAL_LARGE"};
static uint32_t occ_lds(OccKernel k)
{
    if (const char* e = getenv(kOccEnv[k])) return (uint32_t)std::min(32768L, std::max(0L, atol(e)));
    return kOccDefault[k];
}

// Block of a batch position: off[] holds the block offsets; tab[k] is the
// block holding position k << kTabShift (tab has one entry more than there
// are buckets, the last one nblk - 1), so a lookup searches only between the
// blocks of its own and of the next bucket: usually one or two candidates.
constexpr int kTabShift = 12;
struct BlockMap {
    const uint32_t* off;
    const uint32_t* tab;
    uint32_t nblk;
    __device__ __forceinline__ uint32_t find(uint32_t i) const
    {
        const uint32_t k = i >> kTabShift;
        uint32_t lo = tab[k], hi = tab[k + 1] + 1;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (off[mid] <= i) lo = mid; else hi = mid;
        }
        return lo;
    }
};

// SA entries. Plain: the suffix's global position. Char mode (cb = 4, a batch
// of at most 15 distinct bytes in blocks below 2^28 bytes): the block-local
// position << 4 | the code of the byte before the suffix, so every sort that
// moves a suffix also moves its BWT symbol and the output is a streaming pass
// (the code of a block's first position is unused: the primary row emits
// none). An entry's block is the block of its SA slot.
__device__ __forceinline__ uint32_t sa_pos(uint32_t v, uint32_t cb, uint32_t block_off)
{
    return cb ? (v >> cb) + block_off : v;
}

// Host side of BlockMap::find: tab[k] = last block b with off[b] <= k << kTabShift.
static void build_block_tab(const std::vector<uint32_t>& offs, uint32_t n, std::vector<uint32_t>& tab)
{
    const uint32_t nblk = (uint32_t)offs.size() - 1;
    const uint32_t nb = (n >> kTabShift) + 1;
    tab.resize((size_t)nb + 1);
    uint32_t b = 0;
    for (uint32_t k = 0; k < nb; ++k) {
        const uint64_t p = (uint64_t)k << kTabShift;
        while (b + 1 < nblk && offs[b + 1] <= p) ++b;
        tab[k] = b;
    }
    tab[nb] = nblk - 1;
}

__device__ __forceinline__ int grp_class(uint32_t size)
{
    if (size <= 64) return 31 - __builtin_clz(size - 1);  // 2 -> 0, 3..4 -> 1, ..., 33..64 -> 5
    if (size <= (uint32_t)kWaveMax) return kClsWave;
    return size <= (uint32_t)kMedMax ? kClsMed : kClsLarge;
}

// New unsorted groups are staged in LDS per workgroup and flushed with one
// pool atomic and one counter atomic per size class, so the device counters
// see a few thousand atomics per round rather than one per wave.
constexpr int kStage = 2048;  // a workgroup touches <= 4096 suffixes -> <= 2048 groups
struct StageHdr {
    uint32_t cnt[kClasses];
    uint32_t n;
    uint32_t base;
};
template <int CAP>
struct StageT {
    StageHdr h;
    uint2 e[CAP];  // (start, size)
};
using Stage = StageT<kStage>;

__device__ __forceinline__ void stage_init(StageHdr& H)
{
    if (threadIdx.x < kClasses) H.cnt[threadIdx.x] = 0;
    if (threadIdx.x == 0) H.n = 0;
}

__device__ __forceinline__ void stage_push_e(StageHdr& H, uint2* E, uint32_t start, uint32_t size, uint32_t cap)
{
    atomicAdd(&H.cnt[grp_class(size)], 1u);
    const uint32_t k = atomicAdd(&H.n, 1u);
    if (k < cap) E[k] = make_uint2(start, size);
}

// Whole workgroup; ends with the staged groups in the device pool.
__device__ __forceinline__ void stage_flush_e(StageHdr& H, const uint2* E, const GroupLists& gl, uint32_t cap)
{
    __syncthreads();
    const uint32_t n = H.n < cap ? H.n : cap;
    if (threadIdx.x < kClasses && H.cnt[threadIdx.x])
        atomicAdd(&gl.cnt[threadIdx.x * kCtrStride], H.cnt[threadIdx.x]);
    if (threadIdx.x == 0) {
        if (H.n > cap) *gl.overflow = 1u;
        H.base = n ? atomicAdd(gl.top, n) : 0u;
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) pool_put(gl, H.base + k, E[k].x, E[k].y);
}

template <int CAP>
__device__ __forceinline__ void stage_init(StageT<CAP>& S_) { stage_init(S_.h); }
template <int CAP>
__device__ __forceinline__ void stage_push_t(StageT<CAP>& S_, uint32_t start, uint32_t size, uint32_t cap)
{
    stage_push_e(S_.h, S_.e, start, size, cap);
}
template <int CAP>
__device__ __forceinline__ void stage_flush_t(StageT<CAP>& S_, const GroupLists& gl, uint32_t cap)
{
    stage_flush_e(S_.h, S_.e, gl, cap);
}

// Pool -> list ordered by class: class c starts at the sum of the smaller
// classes' counts. Each workgroup moves one contiguous range of kBucketRange
// pool entries and keeps their order within every class (wave-major, then
// step, then lane, ranked by ballots), so a class list is made of long runs
// of groups emitted by the same workgroup: neighbouring slots, which the
// next round's kernels read together. fill[] (zeroed) hands out the places
// of each range's run within its class.
constexpr uint32_t kBucketRange = 4096;
__global__ __launch_bounds__(256) void k_bucket(const Grp* __restrict__ pool, uint32_t total, uint32_t down,
                                                uint32_t cap, const uint32_t* __restrict__ cnt,
                                                Grp* __restrict__ list, uint32_t* __restrict__ fill)
{
    // entry i of the set (emission order); a downward set ends at pool[cap - 1]
    const Grp* const at0 = down ? pool + cap - 1u : pool;
    const int step = down ? -1 : 1;
    constexpr uint32_t W = 256 / QZ_WAVE, Q = kBucketRange / W;  // 1024 entries per wave
    __shared__ uint32_t co[kClasses];
    __shared__ uint32_t wb[W][kClasses];  // per wave: class counts, then write bases
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    const uint32_t q0 = blockIdx.x * kBucketRange + w * Q;
    if (threadIdx.x == 0) {
        uint32_t acc = 0;
        for (int c = 0; c < kClasses; ++c) {
            co[c] = acc;
            acc += cnt[c * kCtrStride];
        }
    }
    uint32_t run[kClasses];
#pragma unroll
    for (int c = 0; c < kClasses; ++c) run[c] = 0;
    for (uint32_t b = q0; b < q0 + Q && b < total; b += QZ_WAVE) {  // wave-uniform
        const uint32_t i = b + lane;
        const int cls = i < total ? grp_class(at0[step * (int64_t)i].size) : kClasses;
#pragma unroll
        for (int c = 0; c < kClasses; ++c) run[c] += (uint32_t)__popcll(__ballot(cls == c));
    }
    if (lane < (uint32_t)kClasses) {
        uint32_t v = 0;
#pragma unroll
        for (int c = 0; c < kClasses; ++c) v = (lane == (uint32_t)c) ? run[c] : v;
        wb[w][lane] = v;
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)kClasses) {
        const uint32_t c = threadIdx.x;
        uint32_t tot = 0;
        for (uint32_t v = 0; v < W; ++v) tot += wb[v][c];
        uint32_t base = tot ? atomicAdd(&fill[c * kCtrStride], tot) + co[c] : 0u;
        for (uint32_t v = 0; v < W; ++v) {
            const uint32_t x = wb[v][c];
            wb[v][c] = base;
            base += x;
        }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kClasses; ++c) run[c] = wb[w][c];
    for (uint32_t b = q0; b < q0 + Q && b < total; b += QZ_WAVE) {
        const uint32_t i = b + lane;
        Grp g = {0, 0};
        int cls = kClasses;
        if (i < total) {
            g = at0[step * (int64_t)i];
            cls = grp_class(g.size);
        }
        uint32_t dst = 0;
#pragma unroll
        for (int c = 0; c < kClasses; ++c) {
            const uint64_t m = __ballot(cls == c);
            if (cls == c) dst = run[c] + popc_below(m);
            run[c] += (uint32_t)__popcll(m);
        }
        if (cls < kClasses) list[dst] = g;
    }
}


__device__ __forceinline__ void stage_push(Stage& S, uint32_t start, uint32_t size)
{
    stage_push_e(S.h, S.e, start, size, kStage);
}

// Whole workgroup; ends with the staged groups in the device lists.
__device__ __forceinline__ void stage_flush(Stage& S, const GroupLists& gl) { stage_flush_e(S.h, S.e, gl, kStage); }

// The doubling key of suffix s: 1 + rank of s + d (d = the depth of s's
// group, the symbols all its members share), 0 past the block end.
__device__ __forceinline__ uint32_t round_key(const uint32_t* __restrict__ rank, const BlockMap& bm, uint32_t slot,
                                              uint32_t v, uint32_t cb, uint32_t d)
{
    const uint32_t b = bm.find(slot);
    const uint32_t end = bm.off[b + 1];
    const uint64_t p = (uint64_t)sa_pos(v, cb, bm.off[b]) + d;
    return (p < end) ? rank[(uint32_t)p] + 1u : 0u;
}

// Depth of a new subgroup (>= 2 members) of the group starting at slot
// `start`: the group's depth plus that of the group its members' keys point
// into (key = 1 + that group's start slot). Members with equal keys share
// the group's depth of symbols and then the target group's; a key of 0
// (past the block end) is unique within a group and never forms a subgroup.
__device__ __forceinline__ uint32_t sub_depth(const uint32_t* __restrict__ dep_in, uint32_t start, uint32_t key)
{
    const uint32_t d = dep_in[start];
    const uint32_t t = key ? dep_in[key - 1u] : 0u;
    const uint32_t s = d + t;
    return s < d ? 0xffffffffu : s;  // saturate (never reached: depth < 2^31)
}

__global__ __launch_bounds__(kThreads) void k_byte_hist(const uint8_t* __restrict__ text, uint32_t n,
                                                        uint32_t* __restrict__ hist)
{
    // R copies of each wave's counters, a value's copies in adjacent words,
    // chosen by lane: 16 counts per lane per load, from few distinct bytes
    // (DNA), would otherwise queue on a handful of words
    constexpr uint32_t R = 4;
    __shared__ uint32_t h[4][256 * R];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, rep = tid & (R - 1);
    for (uint32_t j = tid; j < 256 * R; j += kThreads)
        for (int w = 0; w < 4; ++w) h[w][j] = 0;
    __syncthreads();
    // the 16-byte aligned middle in 16-byte loads; the unaligned head and the
    // tail byte by byte (device-resident callers may pass any offset)
    const uint32_t head = (uint32_t)min((uint64_t)((16u - ((uintptr_t)text & 15u)) & 15u), (uint64_t)n);
    const uint32_t nv = (n - head) / 16u;
    const uint32_t tail = head + nv * 16u;
    const uint32_t g = blockIdx.x * kThreads + tid, gs = gridDim.x * kThreads;
    for (uint32_t i = g; i < head; i += gs) atomicAdd(&h[wave][text[i] * R + rep], 1u);
    for (uint32_t i = tail + g; i < n; i += gs) atomicAdd(&h[wave][text[i] * R + rep], 1u);
    const uint4* v = reinterpret_cast<const uint4*>(text + head);
    for (uint32_t i = g; i < nv; i += gs) {
        const uint4 x = v[i];
        const uint32_t w4[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int b = 0; b < 4; ++b) atomicAdd(&h[wave][((w4[k] >> (8 * b)) & 255u) * R + rep], 1u);
    }
    __syncthreads();
    uint32_t t = 0;
    for (int w = 0; w < 4; ++w)
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) t += h[w][tid * R + r];
    if (t) atomicAdd(&hist[tid], t);
}

// Round-0 keys. Packed: the next k0 symbol codes in radix sigma+1. Run
// length (rle): nseg segments of segbits bits, most significant first, each
// one run (or its first cap symbols) as (code, class, value) in (sb, 2, lb)
// bits: class 0 = the run is followed by a smaller symbol (or the block
// end), value = length; 1 = the run goes on past cap symbols, value = cap;
// 2 = followed by a larger symbol, value = cap + 1 - length. A run of c
// followed by a smaller symbol sorts below any longer run of c, a run
// followed by a larger one above it, so these keys order suffixes as their
// text does; equal keys share the sum of their lengths (the key's depth).
// The choice is per block (blocks are never compared in round 0): a block
// whose mean run is at least kRleMinRun symbols takes run-length keys; when
// any block does, every group of the batch carries its own depth.
struct PackParams {
    uint16_t code[256];  // symbol -> 1..sigma (sigma may be 256)
    uint32_t radix;      // sigma + 1
    uint32_t k0;         // symbols per 32-bit key (packed keys)
    uint32_t rle;        // 1: some block has run-length keys (per-group depths)
    uint32_t cb;         // SA entries carry this many bits of preceding-symbol code (0 or 4)
    uint32_t sb, lb, nseg, segbits, cap;
    int32_t force;       // -1 per block by its runs, 0 no block, 1 every block
};
constexpr uint32_t kRleMinRun = 3;

// Whether block b takes run-length keys: brun[b] run boundaries among the
// brun[nblk + b] sampled positions (k_block_runs); a block with fewer than
// 64 samples keeps packed keys.
__device__ __forceinline__ bool block_rle(const PackParams& pp, const BlockMap& bm, const uint32_t* __restrict__ brun,
                                          uint32_t b)
{
    if (pp.force >= 0) return pp.force != 0;
    const uint32_t seen = brun[bm.nblk + b];
    return seen >= 64u && (uint64_t)seen >= (uint64_t)kRleMinRun * ((uint64_t)brun[b] + 1u);
}

// Symbols a round-0 key covers: k0 for packed keys, the sum of the run
// lengths for run-length keys.
__device__ __forceinline__ uint32_t key_depth(uint32_t key, const PackParams& pp, bool rle)
{
    if (!rle) return pp.k0;
    uint32_t d = 0;
    const uint32_t vmask = (1u << pp.lb) - 1u;
    for (uint32_t q = 0; q < pp.nseg; ++q) {
        const uint32_t part = (key >> (32u - (q + 1u) * pp.segbits)) & ((1u << pp.segbits) - 1u);
        if ((part >> (pp.lb + 2u)) == 0u) break;  // past the block end
        const uint32_t cls = (part >> pp.lb) & 3u, v = part & vmask;
        d += (cls == 2u) ? pp.cap + 1u - v : v;
    }
    return d;
}

// Packed prefix key of every position: the next k0 symbol codes in radix
// sigma+1, code 0 past the block end. Each thread builds 16 consecutive keys
// with a sliding window (drop the leading code, append one), restarting at
// block starts, and stores them as four 16-byte vectors.
constexpr int kKeyItems = 16;
constexpr uint32_t kRleLook = 512;  // run-length keys read at most 511 positions ahead

// A thread's kKeyItems consecutive keys, as 16-byte vectors when whole.
__device__ __forceinline__ void store_keys(uint32_t* __restrict__ keys, uint32_t i0, uint32_t n,
                                           const uint32_t (&out)[kKeyItems])
{
    if (i0 + kKeyItems <= n) {
        uint4* dst = reinterpret_cast<uint4*>(keys + i0);
#pragma unroll
        for (int v = 0; v < kKeyItems / 4; ++v)
            dst[v] = make_uint4(out[4 * v], out[4 * v + 1], out[4 * v + 2], out[4 * v + 3]);
    } else {
#pragma unroll
        for (int q = 0; q < kKeyItems; ++q)
            if (i0 + q < n) keys[i0 + q] = out[q];
    }
}
// The SA entry of position i of the block starting at bo; prev is the byte
// before i (unused at a block start).
__device__ __forceinline__ uint32_t sa_entry(uint32_t i, uint32_t bo, uint32_t cb, const uint16_t* code, uint8_t prev)
{
    return cb ? ((i - bo) << cb) | (i > bo ? (uint32_t)code[prev] : 0u) : i;
}

__global__ __launch_bounds__(kThreads) void k_round0_keys(const uint8_t* __restrict__ text, uint32_t n,
                                                          BlockMap bm,
                                                          const PackParams* __restrict__ ppd,
                                                          const uint32_t* __restrict__ brun,
                                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals)
{
    // the workgroup's 4096 positions and the 511 a run-length key may look
    // ahead: W window bytes, W remaining run lengths, W block-start flags;
    // the packed path reuses the same LDS for its 4096 keys
    constexpr uint32_t W = kThreads * kKeyItems + kRleLook;
    // packed path: 4096 keys, then the 4096 positions and the 31 (k0 <= 32)
    // that their keys read ahead
    constexpr uint32_t KB = kThreads * kKeyItems * 4, WP = kThreads * kKeyItems + 32;
    constexpr uint32_t kSmem = 4 * W > KB + WP ? 4 * W : KB + WP;
    static_assert(W % 16 == 0 && WP % 16 == 0 && KB % 16 == 0, "LDS carve");
    __shared__ uint16_t code[256];
    __shared__ __align__(16) uint8_t smem[kSmem];
    code[threadIdx.x] = ppd->code[threadIdx.x];
    const uint32_t radix = ppd->radix, k0 = ppd->k0, cb = ppd->cb;
    __syncthreads();
    const uint32_t i0 = (blockIdx.x * kThreads + threadIdx.x) * kKeyItems;
    if (ppd->rle) {  // uniform: the whole grid takes this branch or none
        // window entries with each entry's remaining run length rem
        // (capped at cap + 1; runs end at block ends)
        constexpr uint32_t C = (W + kThreads - 1) / kThreads;  // entries per thread
        uint8_t* win = smem;
        uint16_t* rem = reinterpret_cast<uint16_t*>(smem + W);
        uint8_t* bstart = smem + 3 * W;  // 1: a block starts here
        const PackParams& pp = *ppd;
        const uint32_t base = blockIdx.x * kThreads * kKeyItems;
        for (uint32_t k = threadIdx.x; k < W; k += kThreads) {
            win[k] = base + k < n ? text[base + k] : 0;
            bstart[k] = 0;
        }
        __syncthreads();
        {
            // block starts inside the window (blocks of any size)
            const uint32_t b0 = bm.find(base < n ? base : n - 1);
            for (uint32_t b = b0 + 1 + threadIdx.x; b < bm.nblk; b += kThreads) {
                const uint32_t o = bm.off[b];
                if (o >= base + W || o >= n) break;
                bstart[o - base] = 1;
            }
        }
        __syncthreads();
        {
            // thread t: entries [t*C, t*C + C): the last one by a bounded
            // forward scan, the others backwards from it
            const uint32_t k0c = threadIdx.x * C;
            const uint32_t k1c = k0c + C < W ? k0c + C : W;
            if (k0c < k1c) {
                uint32_t k = k1c - 1, r = 1;
                while (r <= pp.cap && k + r < W && !bstart[k + r] && base + k + r < n && win[k + r] == win[k]) ++r;
                rem[k] = (uint16_t)r;
                while (k > k0c) {
                    --k;
                    const bool same = base + k + 1 < n && !bstart[k + 1] && win[k + 1] == win[k];
                    r = same ? (r <= pp.cap ? r + 1 : r) : 1u;
                    rem[k] = (uint16_t)r;
                }
            }
        }
        __syncthreads();
        if (i0 >= n) return;
        uint32_t end = 0, bo = 0, out[kKeyItems], vo[kKeyItems];
        bool rleb = false;
#pragma unroll
        for (int q = 0; q < kKeyItems; ++q) {
            const uint32_t i = i0 + q;
            out[q] = 0;
            vo[q] = 0;
            if (i < n) {
                if (i >= end) {
                    const uint32_t b = bm.find(i);
                    bo = bm.off[b];
                    end = bm.off[b + 1];
                    rleb = block_rle(pp, bm, brun, b);
                }
                vo[q] = sa_entry(i, bo, cb, code, i > bo ? (i - 1 >= base ? win[i - 1 - base] : text[i - 1]) : 0);
                if (!rleb) {  // a packed key from the window (k0 <= 32 < kRleLook)
                    uint32_t key = 0;
                    for (uint32_t j = 0; j < k0; ++j) {
                        const uint32_t p = i + j;
                        key = key * radix + ((p < end) ? code[win[p - base]] : 0u);
                    }
                    out[q] = key;
                    continue;
                }
                uint32_t key = 0, p = i;
                for (uint32_t g = 0; g < pp.nseg; ++g) {
                    uint32_t part = 0;
                    if (p < end) {
                        const uint32_t r = rem[p - base];
                        const uint32_t c = code[win[p - base]];
                        uint32_t cls, v, len;
                        if (r > pp.cap) {
                            cls = 1u;
                            v = len = pp.cap;
                        } else {
                            const uint32_t nx = (p + r < end) ? code[win[p + r - base]] : 0u;
                            cls = nx < c ? 0u : 2u;
                            v = nx < c ? r : pp.cap + 1u - r;
                            len = r;
                        }
                        part = (c << (pp.lb + 2u)) | (cls << pp.lb) | v;
                        p += len;
                    }
                    key = (key << pp.segbits) | part;
                }
                out[q] = key << (32u - pp.nseg * pp.segbits);
            }
        }
        store_keys(keys, i0, n, out);
        if (vals) store_keys(vals, i0, n, vo);
        return;
    }
    // packed keys: the workgroup stages its text window in LDS (16-byte
    // loads when the text is aligned), each thread slides its key over 16
    // positions, the keys go to LDS, and the workgroup stores them as
    // contiguous 16-byte vectors
    const uint32_t base = blockIdx.x * kThreads * kKeyItems;
    if (base >= n) return;  // uniform over the workgroup
    uint32_t* kl = reinterpret_cast<uint32_t*>(smem);
    uint8_t* win = smem + KB;
    {
        const uint8_t* tb = text + base;
        const bool al = ((uintptr_t)tb & 15u) == 0;
        for (uint32_t v = threadIdx.x; v < WP / 16u; v += kThreads) {
            const uint32_t p = v * 16u;
            if (al && base + p + 16u <= n) {
                *reinterpret_cast<uint4*>(win + p) = *reinterpret_cast<const uint4*>(tb + p);
            } else {
                for (uint32_t b = 0; b < 16u; ++b) win[p + b] = base + p + b < n ? tb[p + b] : 0;
            }
        }
    }
    __syncthreads();
    uint32_t vo[kKeyItems];
    if (i0 < n) {
        uint32_t top = 1;  // weight of the leading code
        for (uint32_t j = 1; j < k0; ++j) top *= radix;
        uint32_t end = 0, bo = 0, key = 0;
        uint32_t out[kKeyItems];
#pragma unroll
        for (int q = 0; q < kKeyItems; ++q) {
            const uint32_t i = i0 + q;
            vo[q] = 0;
            if (i < n) {
                if (i >= end) {  // first position or a new block: full key
                    const uint32_t b = bm.find(i);
                    bo = bm.off[b];
                    end = bm.off[b + 1];
                    key = 0;
                    for (uint32_t j = 0; j < k0; ++j) {
                        const uint32_t p = i + j;
                        key = key * radix + ((p < end) ? code[win[p - base]] : 0u);
                    }
                } else {
                    const uint32_t lead = code[win[i - 1 - base]];
                    const uint32_t p = i + k0 - 1;
                    key = (key - lead * top) * radix + ((p < end) ? code[win[p - base]] : 0u);
                }
                vo[q] = sa_entry(i, bo, cb, code, i > bo ? (i - 1 >= base ? win[i - 1 - base] : text[i - 1]) : 0);
            }
            out[q] = key;
        }
        uint4* dst = reinterpret_cast<uint4*>(kl + threadIdx.x * kKeyItems);
#pragma unroll
        for (int v = 0; v < kKeyItems / 4; ++v)
            dst[v] = make_uint4(out[4 * v], out[4 * v + 1], out[4 * v + 2], out[4 * v + 3]);
    }
    // keys, then SA entries, each staged in LDS and stored as contiguous
    // 16-byte vectors
    const uint32_t cnt = min(n - base, (uint32_t)(kThreads * kKeyItems));
    const uint4* src = reinterpret_cast<const uint4*>(kl);
    for (int pass = 0; pass < (vals ? 2 : 1); ++pass) {  // vals == nullptr: keys only
        uint32_t* o = pass ? vals : keys;
        if (pass) {
            __syncthreads();  // the keys are out of LDS
            if (i0 < n) {
                uint4* dst = reinterpret_cast<uint4*>(kl + threadIdx.x * kKeyItems);
#pragma unroll
                for (int v = 0; v < kKeyItems / 4; ++v)
                    dst[v] = make_uint4(vo[4 * v], vo[4 * v + 1], vo[4 * v + 2], vo[4 * v + 3]);
            }
        }
        __syncthreads();
        uint4* ov = reinterpret_cast<uint4*>(o + base);
        for (uint32_t v = threadIdx.x; v < cnt / 4u; v += kThreads) ov[v] = src[v];
        for (uint32_t i = (cnt & ~3u) + threadIdx.x; i < cnt; i += kThreads) o[base + i] = kl[i];
    }
}

// Symbol packing from the byte histogram (one workgroup of 256): present
// symbols get codes 1..sigma in byte order, 0 means past the block end, and
// k0 is the largest k with (sigma+1)^k <= 2^32.
// Sampled run boundaries per block: one thread per kRunSample-th position
// counts, in the aligned 16-byte window holding it, the neighbouring pairs
// inside the block that differ (brun[b]) and the pairs it looked at
// (brun[nblk + b]);
// summed per wave when the wave lies in one block. The key mode only needs
// the mean run length, which a 1/64 sample estimates well.
constexpr uint32_t kRunSample = 1024;
__global__ __launch_bounds__(kThreads) void k_block_runs(const uint8_t* __restrict__ text, uint32_t n, BlockMap bm,
                                                         uint32_t* __restrict__ brun)
{
    const uint32_t i0 = (blockIdx.x * kThreads + threadIdx.x) * kRunSample;
    const bool live = i0 < n;
    const uint32_t b = live ? bm.find(i0) : 0xffffffffu;
    uint32_t cnt = 0, seen = 0;
    if (live) {
        // one 16-byte load: the aligned window that holds position i0 (it
        // holds a byte of the text, so it lies in the text's pages); the
        // neighbouring pairs of the window inside i0's block
        const uint32_t bs = bm.off[b], be = bm.off[b + 1];
        const uint32_t mis = (uint32_t)((uintptr_t)(text + i0) & 15u);
        const uint4 v = *reinterpret_cast<const uint4*>(text + i0 - mis);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        const int64_t p0 = (int64_t)i0 - mis;
        uint32_t prev = 0;
        bool have = false;
#pragma unroll
        for (uint32_t q = 0; q < 16; ++q) {
            const int64_t p = p0 + q;
            if (p < (int64_t)bs || p >= (int64_t)be) continue;
            const uint32_t c = (w4[q >> 2] >> (8u * (q & 3u))) & 255u;
            if (have) {
                cnt += c != prev ? 1u : 0u;
                ++seen;
            }
            prev = c;
            have = true;
        }
    }
    const uint32_t b0 = __shfl(b, 0, QZ_WAVE);
    if (!__ballot(b != b0)) {  // the whole wave samples one block
        for (int d = 1; d < QZ_WAVE; d <<= 1) {
            cnt += __shfl_xor(cnt, d, QZ_WAVE);
            seen += __shfl_xor(seen, d, QZ_WAVE);
        }
        if (lane_id() == 0 && b0 != 0xffffffffu) {
            if (cnt) atomicAdd(&brun[b0], cnt);
            if (seen) atomicAdd(&brun[bm.nblk + b0], seen);
        }
    } else if (live) {
        if (cnt) atomicAdd(&brun[b], cnt);
        if (seen) atomicAdd(&brun[bm.nblk + b], seen);
    }
}

// Symbol packing and the key mode (force: -1 per block by its runs, 0 never
// run-length keys, 1 always).
__global__ __launch_bounds__(kThreads) void k_make_pack(const uint32_t* __restrict__ hist, PackParams* __restrict__ pp,
                                                        int force, BlockMap bm, const uint32_t* __restrict__ brun,
                                                        uint32_t maxb_bits)
{
    __shared__ uint32_t any_rle;
    if (threadIdx.x == 0) any_rle = 0;
    __syncthreads();
    if (force < 0) {
        for (uint32_t b = threadIdx.x; b < bm.nblk; b += kThreads) {
            const uint32_t seen = brun[bm.nblk + b];
            if (seen >= 64u && (uint64_t)seen >= (uint64_t)kRleMinRun * ((uint64_t)brun[b] + 1u)) any_rle = 1;
        }
    }
    __syncthreads();
    __shared__ uint32_t wsum[kThreads / QZ_WAVE];
    const uint32_t c = threadIdx.x, lane = lane_id(), w = c / QZ_WAVE;
    const uint32_t present = hist[c] ? 1u : 0u;
    uint32_t incl = present;
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, QZ_WAVE);
        if (lane >= (uint32_t)d) incl += o;
    }
    if (lane == QZ_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    for (uint32_t i = 0; i < w; ++i) incl += wsum[i];
    pp->code[c] = present ? (uint16_t)incl : (uint16_t)0;
    if (c == kThreads - 1) {
        const uint64_t radix = (uint64_t)incl + 1;
        uint64_t v = 1;
        uint32_t k = 0;
        while (k < 32 && v * radix <= (1ull << 32)) {
            v *= radix;
            ++k;
        }
        pp->radix = (uint32_t)radix;
        pp->k0 = k ? k : 1;
        // char mode: codes 1..15 in 4 bits, or 1..63 in 6, beside
        // block-local positions that fit the other bits of the entry
        const uint32_t cbits = radix <= 16 ? 4u : radix <= 64 ? 6u : 0u;
        pp->cb = (cbits && maxb_bits + cbits <= 32u) ? cbits : 0u;
        const bool rle = force >= 0 ? force != 0 : any_rle != 0;
        uint32_t sb = 0;
        while ((1ull << sb) < radix) ++sb;  // codes 0..sigma
        const uint32_t nseg = sb <= 5 ? 3u : 2u;
        const uint32_t segbits = 32u / nseg;  // 10 or 16
        const uint32_t lb = segbits - 2u - sb;
        pp->rle = rle ? 1u : 0u;
        pp->sb = sb;
        pp->lb = lb;
        pp->nseg = nseg;
        pp->segbits = segbits;
        pp->cap = (1u << lb) - 1u;
        pp->force = force;
    }
}

__global__ __launch_bounds__(kThreads) void k_zero_u32(uint32_t* __restrict__ p, uint32_t n)
{
    for (uint32_t i = threadIdx.x; i < n; i += kThreads) p[i] = 0;
}

// Round-0 group heads, tile by tile over the sorted radix tiles (4096 slots,
// never spanning two blocks). A slot is a head at its block's start or where
// the key changes. lh[t] = the last head slot of tile t (0 if none: max-scan
// identity, slots being >= 0).
__device__ __forceinline__ bool slot_head(const uint32_t* __restrict__ keys, const TileDesc& td, uint32_t i)
{
    const uint32_t j = td.start + i;
    return (i == 0 && td.tlocal == 0) || keys[j] != keys[j - 1];
}

__global__ __launch_bounds__(kThreads) void k_tile_lasthead(const uint32_t* __restrict__ keys,
                                                            const TileDesc* __restrict__ tiles,
                                                            uint32_t* __restrict__ lh)
{
    // backwards in slices of 256 slots, stopping at the first slice that
    // holds a head (keys that change often: one slice)
    __shared__ uint32_t wmax[kThreads / QZ_WAVE];
    const TileDesc td = tiles[blockIdx.x];
    uint32_t r = 0;
    for (uint32_t hi = td.len; hi > 0;) {
        const uint32_t lo = hi > kThreads ? hi - kThreads : 0;
        const uint32_t i = lo + threadIdx.x;
        const bool h = i < hi && slot_head(keys, td, i);
        if (__syncthreads_or(h)) {  // the same for every thread
            uint32_t m = h ? td.start + i : 0u;
            for (int d = 1; d < QZ_WAVE; d <<= 1) {
                const uint32_t o = __shfl_xor(m, d, QZ_WAVE);
                m = o > m ? o : m;
            }
            if (lane_id() == 0) wmax[threadIdx.x / QZ_WAVE] = m;
            __syncthreads();
            for (int w = 0; w < kThreads / QZ_WAVE; ++w) r = wmax[w] > r ? wmax[w] : r;
            break;
        }
        hi = lo;
    }
    if (threadIdx.x == 0) lh[blockIdx.x] = r;
}

// Whether round 0 finishes with the first doubling round fused in
// (k_final0f): the host asks for it (fuse != 0) and the batch has packed
// keys, a choice made on the device.
__device__ __forceinline__ bool fuse_round1(const PackParams& pp, uint32_t fuse)
{
    return fuse != 0u && pp.rle == 0u;
}

// Whether the fused round reads its keys from the 4-bit text (k_pack_nib):
// codes of at most 4 bits, packed keys of at most kNibKeyMax symbols (their
// codes then lie in two aligned 64-bit words).
constexpr uint32_t kNibKeyMax = 17;
__device__ __forceinline__ bool nib_keys(const PackParams& pp)
{
    return pp.radix <= 16u && pp.rle == 0u && pp.k0 <= kNibKeyMax;
}

// Round-0 finish, one tile per workgroup: SA, ranks and the unsorted groups.
// pre[t] = the last head slot before tile t (exclusive max-scan of lh). Wave
// w takes the tile's quarter [1024w, 1024w+1024) in 16 coalesced steps and
// carries the running group start with a wave max-scan; its starting value
// comes from pre[t] and the earlier quarters' last heads.
__global__ __launch_bounds__(kThreads) void k_final0(const uint32_t* __restrict__ keys,
                                                     const uint32_t* __restrict__ vals,
                                                     const TileDesc* __restrict__ tiles,
                                                     const uint32_t* __restrict__ pre, uint32_t* __restrict__ sa,
                                                     uint32_t* __restrict__ rank, GroupLists gl,
                                                     const PackParams* __restrict__ ppd, BlockMap bm, uint32_t fuse)
{
    if (fuse_round1(*ppd, fuse)) return;  // uniform: k_final0f finishes this batch
    constexpr uint32_t Q = kSortTile / (kThreads / QZ_WAVE);  // 1024 slots per wave
    __shared__ Stage S;
    __shared__ uint32_t qlast[kThreads / QZ_WAVE];
    stage_init(S);
    const TileDesc td = tiles[blockIdx.x];
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    const uint32_t qb = w * Q;
    const bool last_tile = td.tlocal + 1 == td.ntiles;
    // All of a lane's keys, neighbouring keys and suffixes first, clamped as
    // in k_final0f.
    constexpr uint32_t SQ = Q / QZ_WAVE;
    const uint32_t lastj = td.start + td.len - 1u;
    uint32_t kc[SQ], kp[SQ], kn[SQ], vs[SQ];
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t j = min(td.start + qb + st * QZ_WAVE + lane, lastj);
        kc[st] = keys[j];
        kp[st] = keys[j ? j - 1u : 0u];
        kn[st] = keys[last_tile ? min(j + 1u, lastj) : j + 1u];
        vs[st] = vals[j];
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t i = qb + st * QZ_WAVE + lane;
        if (i < td.len && ((i == 0 && td.tlocal == 0) || kc[st] != kp[st])) m = td.start + i;
    }
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_xor(m, d, QZ_WAVE);
        m = o > m ? o : m;
    }
    if (lane == 0) qlast[w] = m;
    __syncthreads();
    uint32_t carry = pre[blockIdx.x];
    for (uint32_t v = 0; v < w; ++v) carry = qlast[v] > carry ? qlast[v] : carry;
    const uint32_t cb = ppd->cb, bo = bm.off[td.block];
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t b = qb + st * QZ_WAVE;
        if (b >= td.len) break;  // wave-uniform
        const uint32_t i = b + lane;
        const bool valid = i < td.len;
        const uint32_t j = td.start + i;
        const uint32_t k = kc[st];
        const bool head = valid && ((i == 0 && td.tlocal == 0) || k != kp[st]);
        uint32_t g = head ? j : 0u;
        for (int d = 1; d < QZ_WAVE; d <<= 1) {
            const uint32_t o = __shfl_up(g, d, QZ_WAVE);
            if (lane >= (uint32_t)d && o > g) g = o;
        }
        if (carry > g) g = carry;
        carry = __shfl(g, QZ_WAVE - 1, QZ_WAVE);
        if (valid) {
            const uint32_t s = vs[st];
            sa[j] = s;
            rank[sa_pos(s, cb, bo)] = g;
            const bool end = (last_tile && i + 1 == td.len) || kn[st] != k;
            if (end && j > g) stage_push(S, g, j - g + 1);
        }
    }
    stage_flush(S, gl);
}

// Run-length keys: the depth of every group round 0 left (set A, pool
// entries [0, total)), from its round-0 key, which the sorted keys still
// hold at the group's start slot.
__global__ __launch_bounds__(kThreads) void k_init_depths(const Grp* __restrict__ pool, uint32_t total,
                                                          const uint32_t* __restrict__ keys,
                                                          const PackParams* __restrict__ ppd, BlockMap bm,
                                                          const uint32_t* __restrict__ brun, uint32_t* __restrict__ dep)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i >= total) return;
    const uint32_t g = pool[i].start;
    dep[g] = key_depth(keys[g], *ppd, block_rle(*ppd, bm, brun, bm.find(g)));
}

// ---- round 0 finish with the first doubling round fused in ----------------
// With packed keys, the first doubling round sorts each round-0 group by
// key2 = rank[s + k0] + 1. The round-0 key of position s + k0 (the packed
// code of its k0 symbols, 0 past the block end) orders those suffixes
// exactly as their round-0 ranks do (equal codes <=> the same round-0
// group), so for a group that lies inside its radix tile and has at most 64
// members, the finishing workgroup sorts it right away by that key, read
// from the positional round-0 keys, which the radix sort leaves intact:
// the group's members get the SA slots and ranks of the first doubling
// round directly, and its subgroups (2 k0 symbols deep) go to set B, which
// the second doubling round reads.
// Every other group of size >= 2 goes to set A as in k_final0; the first
// doubling round then sorts set A alone (its keys may point into groups
// already refined here, which only orders them further: ranks stay the
// start slots of groups in sorted order) and appends to set B.
//
// Phases: (1) group heads as in k_final0; slots of singletons and of groups
// that start before the tile are final at once, the others wait in LDS; the
// ends of in-tile groups of <= 64 members list them as fused, the ends of
// other groups stage them for set A. (2) every waiting slot of an unfused
// group gets its SA entry and rank; every slot of a fused group its code.
// (3) waves sort the fused groups by code, by size class as k_sort_small
// does, and stage their subgroups for set B.
constexpr uint32_t kFuseMax = 64;
constexpr uint32_t kWaitNone = 0xffffffffu;  // slot already final

// Two-ended stage: set A entries from e[0] up, set B entries from e[kStage-1]
// down (a tile's groups and subgroups are disjoint, so <= kStage in all).
struct Stage2 {
    uint32_t cnt[2][kClasses];
    uint32_t n[2];
    uint32_t base[2];
    uint2 e[kStage];
};

__device__ __forceinline__ void stage2_push(Stage2& S, uint32_t set, uint32_t start, uint32_t size)
{
    atomicAdd(&S.cnt[set][grp_class(size)], 1u);
    const uint32_t k = atomicAdd(&S.n[set], 1u);
    if (k < (uint32_t)kStage) S.e[set ? (uint32_t)kStage - 1u - k : k] = make_uint2(start, size);
}

__device__ __forceinline__ void stage2_flush(Stage2& S, const GroupLists& glA, const GroupLists& glB)
{
    __syncthreads();
    const uint32_t t = threadIdx.x;
    if (t < (uint32_t)kClasses) {
        if (S.cnt[0][t]) atomicAdd(&glA.cnt[t * kCtrStride], S.cnt[0][t]);
        if (S.cnt[1][t]) atomicAdd(&glB.cnt[t * kCtrStride], S.cnt[1][t]);
    }
    if (t == 0) {
        if (S.n[0] + S.n[1] > (uint32_t)kStage) *glA.overflow = 1u;
        S.base[0] = S.n[0] ? atomicAdd(glA.top, S.n[0]) : 0u;
        S.base[1] = S.n[1] ? atomicAdd(glB.top, S.n[1]) : 0u;
    }
    __syncthreads();
    const uint32_t na = min(S.n[0], (uint32_t)kStage), nb = min(S.n[1], (uint32_t)kStage);
    for (uint32_t k = t; k < na; k += blockDim.x) pool_put(glA, S.base[0] + k, S.e[k].x, S.e[k].y);
    for (uint32_t k = t; k < nb; k += blockDim.x) {
        const uint2 e = S.e[(uint32_t)kStage - 1u - k];
        pool_put(glB, S.base[1] + k, e.x, e.y);
    }
}

// T threads per tile; SVL: waiting slots keep their suffix in LDS (else it
// is read again from vals, trading 16 KB of LDS for a coalesced re-read).
// kpos: the round-0 key of every position.
template <uint32_t T, bool SVL>
__global__ __launch_bounds__(T) void k_final0f(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                               const TileDesc* __restrict__ tiles, const uint32_t* __restrict__ pre,
                                               uint32_t* __restrict__ sa, uint32_t* __restrict__ rank,
                                               GroupLists glA, GroupLists glB, const PackParams* __restrict__ ppd,
                                               BlockMap bm, const uint32_t* __restrict__ kpos,
                                               const uint32_t* __restrict__ nib, uint32_t fuse)
{
    if (!fuse_round1(*ppd, fuse)) return;  // uniform: k_final0 finishes this batch
    constexpr uint32_t W = T / QZ_WAVE;
    constexpr uint32_t Q = kSortTile / W;      // slots per wave in phase 1
    constexpr uint32_t PER = kSortTile / T;    // slots per thread in phase 2
    __shared__ uint32_t sv[SVL ? kSortTile : 1];  // waiting slots: the suffix
    __shared__ uint32_t sx[kSortTile];         // waiting slots: tile-local group head, then (fused) the code
    __shared__ uint32_t fl[kSortTile / 32];    // fused groups, by tile-local head
    __shared__ uint32_t fe[kStage];            // fused groups as (head << 8) | size, in any order
    __shared__ uint32_t le[kStage];            // the same by size class
    __shared__ uint32_t fcnt[kSmallClasses], fcur[kSmallClasses], fbase[kSmallClasses + 1], tb[kSmallClasses + 1];
    __shared__ uint32_t nfe;
    __shared__ uint32_t qlast[W];
    __shared__ Stage2 S;
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / QZ_WAVE;
    if (tid < (uint32_t)kClasses) S.cnt[0][tid] = S.cnt[1][tid] = 0u;
    if (tid < 2u) S.n[tid] = 0u;
    if (tid < (uint32_t)kSmallClasses) fcnt[tid] = 0u;
    if (tid == 0) nfe = 0u;
    for (uint32_t k = tid; k < kSortTile / 32; k += T) fl[k] = 0u;
    const TileDesc td = tiles[blockIdx.x];
    const uint32_t qb = w * Q;
    const bool last_tile = td.tlocal + 1 == td.ntiles;
    // Each lane's keys, neighbouring keys and suffixes for all of its steps,
    // loaded before any is used (one wait for the wave, not one per step).
    // Indexes are clamped to the tile's last slot (lanes past it use
    // nothing); j - 1 is clamped at 0 (the head test does not read it for a
    // block's first slot) and j + 1 stays inside a block's last tile (its
    // last slot ends its group without reading it).
    constexpr uint32_t SQ = Q / QZ_WAVE;
    const uint32_t lastj = td.start + td.len - 1u;
    uint32_t kc[SQ], kp[SQ], kn[SQ], vs[SQ];
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t j = min(td.start + qb + st * QZ_WAVE + lane, lastj);
        kc[st] = keys[j];
        kp[st] = keys[j ? j - 1u : 0u];
        kn[st] = keys[last_tile ? min(j + 1u, lastj) : j + 1u];
        vs[st] = vals[j];
    }
    uint32_t m = 0;
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t i = qb + st * QZ_WAVE + lane;
        if (i < td.len && ((i == 0 && td.tlocal == 0) || kc[st] != kp[st])) m = td.start + i;
    }
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_xor(m, d, QZ_WAVE);
        m = o > m ? o : m;
    }
    if (lane == 0) qlast[w] = m;
    __syncthreads();
    uint32_t carry = pre[blockIdx.x];
    for (uint32_t v = 0; v < w; ++v) carry = qlast[v] > carry ? qlast[v] : carry;
    const uint32_t cb = ppd->cb, bo = bm.off[td.block];
    // phase 1
#pragma unroll
    for (uint32_t st = 0; st < SQ; ++st) {
        const uint32_t b = qb + st * QZ_WAVE;
        if (b >= td.len) break;  // wave-uniform
        const uint32_t i = b + lane;
        const bool valid = i < td.len;
        const uint32_t j = td.start + i;
        const uint32_t k = kc[st];
        const bool head = valid && ((i == 0 && td.tlocal == 0) || k != kp[st]);
        uint32_t g = head ? j : 0u;
        for (int d = 1; d < QZ_WAVE; d <<= 1) {
            const uint32_t o = __shfl_up(g, d, QZ_WAVE);
            if (lane >= (uint32_t)d && o > g) g = o;
        }
        if (carry > g) g = carry;
        carry = __shfl(g, QZ_WAVE - 1, QZ_WAVE);
        if (valid) {
            const uint32_t s = vs[st];
            const bool end = (last_tile && i + 1 == td.len) || kn[st] != k;
            if (g < td.start || (head && end)) {  // final now
                sa[j] = s;
                rank[sa_pos(s, cb, bo)] = g;
                sx[i] = kWaitNone;
            } else {
                if (SVL) sv[i] = s;
                sx[i] = g - td.start;
            }
            if (end && j > g) {
                const uint32_t size = j - g + 1;
                if (g >= td.start && size <= kFuseMax) {
                    const uint32_t gl = g - td.start;
                    atomicOr(&fl[gl >> 5], 1u << (gl & 31u));
                    atomicAdd(&fcnt[31 - __builtin_clz(size - 1)], 1u);
                    const uint32_t at = atomicAdd(&nfe, 1u);
                    if (at < (uint32_t)kStage) fe[at] = (gl << 8) | size;
                } else {
                    stage2_push(S, 0, g, size);
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t o = 0, t = 0;
        for (int c = 0; c < kSmallClasses; ++c) {
            fbase[c] = o;
            fcur[c] = o;
            tb[c] = t;
            o += fcnt[c];
            const uint32_t per = (uint32_t)QZ_WAVE >> (c + 1);  // groups per wave task
            t += (fcnt[c] + per - 1) / per;
        }
        fbase[kSmallClasses] = o;
        tb[kSmallClasses] = t;
    }
    // phase 2: waiting slots (fused groups' keys gathered PER slots at a time)
    {
        const uint32_t k0 = ppd->k0;
        const uint32_t bend = bm.off[td.block + 1];
        uint32_t pos[PER];
        bool want[PER];
#pragma unroll
        for (uint32_t q = 0; q < PER; ++q) {
            const uint32_t i = tid + q * T;
            want[q] = false;
            pos[q] = 0;
            if (i < td.len) {
                const uint32_t x = sx[i];
                if (x != kWaitNone) {
                    const uint32_t s = SVL ? sv[i] : vals[td.start + i];
                    if ((fl[x >> 5] >> (x & 31u)) & 1u) {
                        want[q] = true;
                        pos[q] = sa_pos(s, cb, bo) + k0;
                    } else {
                        sa[td.start + i] = s;
                        rank[sa_pos(s, cb, bo)] = td.start + x;
                    }
                }
            }
        }
        // every lane loads (those with nothing to fetch share one address),
        // so all of a thread's gathers are in flight together
        if (nib_keys(*ppd)) {
            // the key from the 4-bit text (an eighth of the positional keys'
            // bytes, so more of it stays in the L2s): the two aligned 64-bit
            // words holding its codes, then the packing of k_round0_keys
            const uint64_t* const nw = reinterpret_cast<const uint64_t*>(nib);
            const uint32_t radix = ppd->radix;
            uint64_t lo[PER], hi[PER];
#pragma unroll
            for (uint32_t q = 0; q < PER; ++q) {
                const uint32_t p = (want[q] && pos[q] < bend) ? pos[q] : td.start;
                lo[q] = nw[p >> 4];
                hi[q] = nw[(p >> 4) + 1u];
            }
#pragma unroll
            for (uint32_t q = 0; q < PER; ++q) {
                if (!want[q]) continue;
                uint32_t key = 0;
                if (pos[q] < bend) {
                    const uint32_t o = pos[q] & 15u, lim = bend - pos[q];
                    for (uint32_t j = 0; j < k0; ++j) {
                        const uint32_t at = o + j;  // < 16 + kNibKeyMax
                        const uint32_t c = (uint32_t)((at < 16u ? lo[q] >> (4u * at) : hi[q] >> (4u * (at - 16u))) & 15u);
                        key = key * radix + (j < lim ? c : 0u);
                    }
                }
                sx[tid + q * T] = key;
            }
        } else {
            uint32_t got[PER];
#pragma unroll
            for (uint32_t q = 0; q < PER; ++q) got[q] = kpos[(want[q] && pos[q] < bend) ? pos[q] : td.start];
#pragma unroll
            for (uint32_t q = 0; q < PER; ++q)
                if (want[q]) sx[tid + q * T] = pos[q] < bend ? got[q] : 0u;
        }
    }
    __syncthreads();  // class bases (thread 0) and codes are in place
    for (uint32_t k = tid; k < min(nfe, (uint32_t)kStage); k += T) {
        const uint32_t e = fe[k];
        const uint32_t at = atomicAdd(&fcur[31 - __builtin_clz((e & 255u) - 1)], 1u);
        le[at] = e;
    }
    __syncthreads();
    // phase 3: fused groups by class, 64 / seg groups per wave task
    for (uint32_t t = w; t < tb[kSmallClasses]; t += W) {  // wave-uniform
        uint32_t c = 0;
        while (c + 1 < (uint32_t)kSmallClasses && t >= tb[c + 1]) ++c;
        const uint32_t seg = 2u << c;
        const uint32_t gi = (t - tb[c]) * ((uint32_t)QZ_WAVE / seg) + lane / seg;
        const uint32_t e = lane & (seg - 1u);
        uint32_t gl = 0, size = 0;
        if (gi < fcnt[c]) {
            const uint32_t ent = le[fbase[c] + gi];
            gl = ent >> 8;
            size = ent & 255u;
        }
        const bool valid = e < size;
        // sort by (key, suffix): pad lanes sort last even past a key of 2^32 - 1
        uint32_t key = valid ? sx[gl + e] : 0xffffffffu;
        uint32_t val = valid ? (SVL ? sv[gl + e] : vals[td.start + gl + e]) : 0xffffffffu;
        for (uint32_t k = 2; k <= seg; k <<= 1) {
            for (uint32_t jj = k >> 1; jj > 0; jj >>= 1) {
                const uint32_t ok = __shfl_xor(key, (int)jj, QZ_WAVE);
                const uint32_t ov = __shfl_xor(val, (int)jj, QZ_WAVE);
                const bool up = (k == seg) || ((lane & k) == 0);
                const bool lower = (lane & jj) == 0;
                const bool less = ok < key || (ok == key && ov < val);
                const bool more = ok > key || (ok == key && ov > val);
                if ((lower == up) ? less : more) {
                    key = ok;
                    val = ov;
                }
            }
        }
        const uint32_t kprev = __shfl_up(key, 1, QZ_WAVE);
        const uint32_t knext = __shfl_down(key, 1, QZ_WAVE);
        const bool head = (e == 0) || kprev != key;
        uint32_t hp = head ? e : 0u;
        for (uint32_t d = 1; d < seg; d <<= 1) {
            const uint32_t o = __shfl_up(hp, d, QZ_WAVE);
            if (e >= d && o > hp) hp = o;
        }
        if (valid) {
            const uint32_t g = td.start + gl;
            sa[g + e] = val;
            rank[sa_pos(val, cb, bo)] = g + hp;
            const bool end = (e + 1 == size) || knext != key;
            if (end && e > hp) stage2_push(S, 1, g + hp, e - hp + 1);
        }
    }
    stage2_flush(S, glA, glB);
}

// ---- small groups (<= 64), one wave holds 64 / 2^(c+1) groups of class c --
struct SmallPlan {
    uint32_t wave_base[kSmallClasses + 1];  // prefix of waves per class
};

__device__ __forceinline__ bool small_lane(const Grp* const* cur, const uint32_t* __restrict__ cnt, const SmallPlan& plan, uint32_t gwave,
                                           int* cls, uint32_t* start, uint32_t* size, uint32_t* e, uint32_t* segw)
{
    int c = 0;
    while (c < kSmallClasses - 1 && gwave >= plan.wave_base[c + 1]) ++c;
    const uint32_t seg = 2u << c;  // lanes per group
    const uint32_t lane = lane_id();
    const uint32_t gi = (gwave - plan.wave_base[c]) * (QZ_WAVE / seg) + lane / seg;
    *cls = c;
    *segw = seg;
    *e = lane & (seg - 1);
    if (gi >= cnt[c * kCtrStride]) return false;
    const Grp g = cur[c][gi];
    *start = g.start;
    *size = g.size;
    return *e < g.size;
}

struct CurLists {
    const Grp* list[kClasses];
    const uint32_t* cnt;
};

constexpr int kGatherTasks = 4;  // 1 and 2 were measured 0.5-1 ms slower per forward on 58 x 25 MiB

// A wave gathers the keys of TASKS consecutive wave tasks. Each load level
// (group descriptors; suffixes, depths and block-table entries; block
// offsets; ranks) is issued for every task before the next level waits on
// it, with indexes clamped to entries that exist, so a wave keeps TASKS
// dependent chains in flight instead of one.
template <int TASKS>
__global__ __launch_bounds__(kThreads) void k_gather_small(CurLists cl, SmallPlan plan,
                                                           const uint32_t* __restrict__ sa,
                                                           const uint32_t* __restrict__ rank,
                                                           BlockMap bm, uint32_t cb, uint32_t h,
                                                           const uint32_t* __restrict__ dep,
                                                           uint32_t* __restrict__ key2)
{
    const uint32_t total = plan.wave_base[kSmallClasses];
    const uint32_t w0 = (blockIdx.x * (kThreads / QZ_WAVE) + threadIdx.x / QZ_WAVE) * TASKS;
    if (w0 >= total) return;  // gather has no barriers
    const uint32_t lane = lane_id();
    uint32_t gi[TASKS], e[TASKS], cnt[TASKS];
    int cls[TASKS];
#pragma unroll
    for (int t = 0; t < TASKS; ++t) {
        const uint32_t gw = min(w0 + (uint32_t)t, total - 1u);  // a class with waves has groups
        int c = 0;
        while (c < kSmallClasses - 1 && gw >= plan.wave_base[c + 1]) ++c;
        const uint32_t seg = 2u << c;
        cls[t] = c;
        gi[t] = (gw - plan.wave_base[c]) * (QZ_WAVE / seg) + lane / seg;
        e[t] = lane & (seg - 1u);
        cnt[t] = cl.cnt[c * kCtrStride];
    }
    Grp g[TASKS];
#pragma unroll
    for (int t = 0; t < TASKS; ++t) g[t] = cl.list[cls[t]][min(gi[t], cnt[t] - 1u)];
    bool valid[TASKS];
    uint32_t slot[TASKS], v[TASKS], d[TASKS], lo[TASKS], hi[TASKS];
#pragma unroll
    for (int t = 0; t < TASKS; ++t) {
        valid[t] = w0 + (uint32_t)t < total && gi[t] < cnt[t] && e[t] < g[t].size;
        slot[t] = g[t].start + (valid[t] ? e[t] : 0u);
        v[t] = sa[slot[t]];
        d[t] = dep ? dep[g[t].start] : h;
        const uint32_t k = slot[t] >> kTabShift;
        lo[t] = bm.tab[k];
        hi[t] = bm.tab[k + 1] + 1u;
    }
    uint32_t ob[TASKS], oe[TASKS];
#pragma unroll
    for (int t = 0; t < TASKS; ++t) {
        while (hi[t] - lo[t] > 1) {  // rare: blocks smaller than a bucket
            const uint32_t mid = (lo[t] + hi[t]) >> 1;
            if (bm.off[mid] <= slot[t]) lo[t] = mid; else hi[t] = mid;
        }
        ob[t] = bm.off[lo[t]];
        oe[t] = bm.off[lo[t] + 1];
    }
    uint32_t p[TASKS], r[TASKS];
#pragma unroll
    for (int t = 0; t < TASKS; ++t) {
        const uint64_t q = (uint64_t)sa_pos(v[t], cb, ob[t]) + d[t];
        p[t] = q < oe[t] ? (uint32_t)q : 0xffffffffu;
        r[t] = rank[p[t] != 0xffffffffu ? p[t] : slot[t]];
    }
#pragma unroll
    for (int t = 0; t < TASKS; ++t)
        if (valid[t]) key2[slot[t]] = p[t] != 0xffffffffu ? r[t] + 1u : 0u;
}

constexpr int kSmallTasks = 16;  // wave tasks per wave in k_sort_small (<= 4096 suffixes per workgroup)

__global__ __launch_bounds__(kThreads) void k_sort_small(CurLists cl, SmallPlan plan, const uint32_t* __restrict__ key2,
                                                         uint32_t* __restrict__ sa, uint32_t* __restrict__ rank,
                                                         GroupLists gl, const uint32_t* __restrict__ dep_in,
                                                         uint32_t* __restrict__ dep_out, BlockMap bm, uint32_t cb)
{
    __shared__ Stage S;
    stage_init(S);
    __syncthreads();
    const uint32_t lane = lane_id();
    const uint32_t wave_g = blockIdx.x * (kThreads / QZ_WAVE) + threadIdx.x / QZ_WAVE;
    for (int t = 0; t < kSmallTasks; ++t) {
        const uint32_t gwave = wave_g * kSmallTasks + t;
        if (gwave >= plan.wave_base[kSmallClasses]) break;  // wave-uniform
        // the task's group (clamped to one that exists: a class with waves
        // has groups), then its keys, suffixes and, in char mode, the block
        // table entries of the rank stores, all issued before any is used
        int c = 0;
        while (c < kSmallClasses - 1 && gwave >= plan.wave_base[c + 1]) ++c;
        const uint32_t seg = 2u << c;
        const uint32_t gi = (gwave - plan.wave_base[c]) * (QZ_WAVE / seg) + lane / seg;
        const uint32_t e = lane & (seg - 1u);
        const uint32_t ncls = cl.cnt[c * kCtrStride];
        const Grp grp = cl.list[c][min(gi, ncls - 1u)];
        const uint32_t start = grp.start, size = grp.size;
        const bool valid = gi < ncls && e < size;
        const uint32_t slot = start + (valid ? e : 0u);
        uint32_t key = key2[slot], val = sa[slot];
        uint32_t blo = 0, bhi = 0;
        if (cb) {  // uniform
            const uint32_t k = start >> kTabShift;
            blo = bm.tab[k];
            bhi = bm.tab[k + 1] + 1u;
        }
        if (!valid) {
            key = kPadKey;
            val = 0;
        }
        // bitonic sort inside aligned segments of `seg` lanes, ascending
        for (uint32_t k = 2; k <= seg; k <<= 1) {
            for (uint32_t j = k >> 1; j > 0; j >>= 1) {
                const uint32_t ok = __shfl_xor(key, (int)j, QZ_WAVE);
                const uint32_t ov = __shfl_xor(val, (int)j, QZ_WAVE);
                const bool up = (k == seg) || ((lane & k) == 0);
                const bool lower = (lane & j) == 0;
                const bool take = (lower == up) ? (ok < key) : (ok > key);
                if (take) {
                    key = ok;
                    val = ov;
                }
            }
        }
        // subgroup heads and their positions within the segment
        const uint32_t kprev = __shfl_up(key, 1, QZ_WAVE);
        const uint32_t knext = __shfl_down(key, 1, QZ_WAVE);
        const bool head = (e == 0) || kprev != key;
        uint32_t hp = head ? e : 0u;
        for (uint32_t d = 1; d < seg; d <<= 1) {
            const uint32_t o = __shfl_up(hp, d, QZ_WAVE);
            if (e >= d && o > hp) hp = o;
        }
        if (valid) {
            sa[start + e] = val;
            if (hp) {  // hp == 0: still the group's start
                uint32_t bo = 0;
                if (cb) {
                    while (bhi - blo > 1) {  // rare: blocks smaller than a bucket
                        const uint32_t mid = (blo + bhi) >> 1;
                        if (bm.off[mid] <= start) blo = mid; else bhi = mid;
                    }
                    bo = bm.off[blo];
                }
                rank[sa_pos(val, cb, bo)] = start + hp;
            }
            const bool end = (e + 1 == size) || knext != key;
            if (end && e > hp) {
                stage_push(S, start + hp, e - hp + 1);
                if (dep_out) dep_out[start + hp] = sub_depth(dep_in, start, key);
            }
        }
        (void)c;
    }
    stage_flush(S, gl);
}

// ---- key gathers of the wave and medium classes (phase A) ----------------
__global__ __launch_bounds__(QZ_WAVE) void k_gather_wave(const Grp* __restrict__ list,
                                                         const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ sa,
                                                         const uint32_t* __restrict__ rank,
                                                         BlockMap bm, uint32_t cb, uint32_t h,
                                                         const uint32_t* __restrict__ dep, uint32_t* __restrict__ key2)
{
    if (blockIdx.x >= *cnt) return;
    const Grp g = list[blockIdx.x];
    const uint32_t d = dep ? dep[g.start] : h;
    for (uint32_t i = threadIdx.x; i < g.size; i += QZ_WAVE) {
        const uint32_t slot = g.start + i;
        key2[slot] = round_key(rank, bm, slot, sa[slot], cb, d);
    }
}

__global__ __launch_bounds__(kThreads) void k_gather_med(const Grp* __restrict__ list, const uint32_t* __restrict__ cnt,
                                                         const uint32_t* __restrict__ sa,
                                                         const uint32_t* __restrict__ rank,
                                                         BlockMap bm, uint32_t cb, uint32_t h,
                                                         const uint32_t* __restrict__ dep, uint32_t* __restrict__ key2)
{
    if (blockIdx.x >= *cnt) return;
    const Grp g = list[blockIdx.x];
    const uint32_t d = dep ? dep[g.start] : h;
    for (uint32_t i = threadIdx.x; i < g.size; i += kThreads) {
        const uint32_t slot = g.start + i;
        key2[slot] = round_key(rank, bm, slot, sa[slot], cb, d);
    }
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
__device__ __forceinline__ uint32_t ceil_div(uint32_t a, uint32_t b) { return (a + b - 1) / b; }

// ---- wave (65 .. 1024) and medium (1025 .. 4096) groups: one workgroup per
// group, stable LSD radix sort in LDS -----------------------------------------
// Keys are ranked relative to the group's minimum, so a group needs
// ceil(bits(max - min) / 9) passes (none when all keys are equal), with the
// digit width spread evenly over them. Wave w owns the contiguous chunk
// [w*C, (w+1)*C) of the current order and ranks 64 keys at a time with one
// ballot per digit bit against wave-private LDS counters, so ranking needs no
// barrier; one scan over (digit, wave) then gives every key its slot.
constexpr int kLdsDigitMax = 9;
constexpr int kLdsBins = 1 << kLdsDigitMax;

template <int WAVES>
__device__ __forceinline__ uint32_t block_excl_sum(uint32_t v, uint32_t* wsum)
{
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    uint32_t incl = v;
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, QZ_WAVE);
        if (lane >= (uint32_t)d) incl += o;
    }
    if (WAVES == 1) return incl - v;
    if (lane == QZ_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    uint32_t pre = 0;
    for (uint32_t i = 0; i < w; ++i) pre += wsum[i];
    return pre + incl - v;
}

template <int WAVES>
__device__ __forceinline__ uint32_t block_excl_max(uint32_t v, uint32_t* wsum)
{
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    uint32_t incl = v;
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, QZ_WAVE);
        if (lane >= (uint32_t)d && o > incl) incl = o;
    }
    uint32_t excl = __shfl_up(incl, 1, QZ_WAVE);
    if (lane == 0) excl = 0;
    if (WAVES == 1) return excl;
    if (lane == QZ_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    for (uint32_t i = 0; i < w; ++i) excl = wsum[i] > excl ? wsum[i] : excl;
    return excl;
}

// Sorts n <= WAVES*64*NB (key, value) pairs of (K, V); returns true when the
// result ends in (K2, V2). Starts and ends with a workgroup barrier.
template <int WAVES, int NB>
__device__ bool lds_radix(uint32_t* K, uint32_t* V, uint32_t* K2, uint32_t* V2, uint32_t* cntr, uint32_t* wsum,
                          uint32_t n, uint32_t kmin, uint32_t bits)
{
    constexpr uint32_t T = WAVES * QZ_WAVE;
    constexpr uint32_t DPT = kLdsBins / T;  // counters per thread in the scan
    if (bits == 0) return false;
    const uint32_t passes = ceil_div(bits, (uint32_t)kLdsDigitMax);
    const uint32_t db = ceil_div(bits, passes);
    const uint32_t dmask = (1u << db) - 1u;
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    const uint32_t chunk = ceil_div(n, T) * QZ_WAVE;
    const uint32_t lo = w * chunk;
    const uint32_t hi = lo + chunk < n ? lo + chunk : n;
    uint32_t* Cw = cntr + w * kLdsBins;
    uint32_t *Ki = K, *Vi = V, *Ko = K2, *Vo = V2;
    for (uint32_t p = 0; p < passes; ++p) {
        const uint32_t sh = p * db;
        for (uint32_t i = lane; i < kLdsBins; i += QZ_WAVE) Cw[i] = 0;
        wave_sync();
        uint32_t kk[NB], rk[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t base = lo + b * QZ_WAVE;
            kk[b] = 0;
            rk[b] = 0;
            if (base < hi) {  // wave-uniform
                const uint32_t e = base + lane;
                const bool valid = e < hi;
                const uint32_t k = valid ? Ki[e] : kmin;
                const uint32_t d = ((k - kmin) >> sh) & dmask;
                uint64_t m = __ballot(valid);
                for (uint32_t j = 0; j < db; ++j) {
                    const bool bit = (d >> j) & 1u;
                    const uint64_t bb = __ballot(bit);
                    m &= bit ? bb : ~bb;
                }
                const uint32_t pre = Cw[d];
                const uint32_t below = popc_below(m);
                const uint32_t tot = (uint32_t)__popcll(m);
                kk[b] = k;
                rk[b] = pre + below;
                if (valid && below + 1 == tot) Cw[d] = pre + tot;
            }
        }
        __syncthreads();
        {
            const uint32_t d0 = threadIdx.x * DPT;
            uint32_t tot = 0;
            for (uint32_t q = 0; q < DPT; ++q)
                for (int v = 0; v < WAVES; ++v) tot += cntr[v * kLdsBins + d0 + q];
            uint32_t run = block_excl_sum<WAVES>(tot, wsum);
            for (uint32_t q = 0; q < DPT; ++q)
                for (int v = 0; v < WAVES; ++v) {
                    const uint32_t x = cntr[v * kLdsBins + d0 + q];
                    cntr[v * kLdsBins + d0 + q] = run;
                    run += x;
                }
        }
        __syncthreads();
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            const uint32_t e = lo + b * QZ_WAVE + lane;
            if (e < hi) {
                const uint32_t d = ((kk[b] - kmin) >> sh) & dmask;
                const uint32_t dst = Cw[d] + rk[b];
                Ko[dst] = kk[b];
                Vo[dst] = Vi[e];
            }
        }
        __syncthreads();
        uint32_t* t = Ki; Ki = Ko; Ko = t;
        t = Vi; Vi = Vo; Vo = t;
    }
    return (passes & 1u) != 0;
}

template <int WAVES>
__global__ __launch_bounds__(WAVES * QZ_WAVE) void k_sort_lds(const Grp* __restrict__ list,
                                                              const uint32_t* __restrict__ cnt,
                                                              const uint32_t* __restrict__ key2,
                                                              uint32_t* __restrict__ sa, uint32_t* __restrict__ rank,
                                                              GroupLists gl, const uint32_t* __restrict__ dep_in,
                                                              uint32_t* __restrict__ dep_out, BlockMap bm, uint32_t cb)
{
    constexpr uint32_t T = WAVES * QZ_WAVE;
    constexpr uint32_t MAXN = WAVES == 1 ? kWaveMax : kMedMax;
    constexpr int NB = MAXN / T;
    static_assert(kLdsBins % T == 0, "scan layout");
    __shared__ uint32_t K[MAXN], V[MAXN], K2[MAXN], V2[MAXN];
    __shared__ uint32_t cntr[WAVES * kLdsBins];
    __shared__ uint32_t wsum[2 * WAVES];
    __shared__ StageHdr S;
    if (blockIdx.x >= *cnt) return;  // block-uniform
    const Grp g = list[blockIdx.x];
    const uint32_t tid = threadIdx.x, lane = lane_id(), w = tid / QZ_WAVE;
    stage_init(S);
    uint32_t kmin = 0xffffffffu, kmax = 0;
    for (uint32_t i = tid; i < g.size; i += T) {
        const uint32_t k = key2[g.start + i];
        K[i] = k;
        V[i] = sa[g.start + i];
        kmin = k < kmin ? k : kmin;
        kmax = k > kmax ? k : kmax;
    }
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t a = __shfl_xor(kmin, d, QZ_WAVE), b = __shfl_xor(kmax, d, QZ_WAVE);
        kmin = a < kmin ? a : kmin;
        kmax = b > kmax ? b : kmax;
    }
    if (lane == 0) {
        wsum[w] = kmin;
        wsum[WAVES + w] = kmax;
    }
    __syncthreads();
    for (int v = 0; v < WAVES; ++v) {
        kmin = wsum[v] < kmin ? wsum[v] : kmin;
        kmax = wsum[WAVES + v] > kmax ? wsum[WAVES + v] : kmax;
    }
    __syncthreads();
    const uint32_t span = kmax - kmin;
    const uint32_t bits = span ? 32u - (uint32_t)__builtin_clz(span) : 0u;
    const bool flip = lds_radix<WAVES, NB>(K, V, K2, V2, cntr, wsum, g.size, kmin, bits);
    const uint32_t* Kf = flip ? K2 : K;
    const uint32_t* Vf = flip ? V2 : V;
    uint32_t* H = flip ? K : K2;
    uint2* E = reinterpret_cast<uint2*>(flip ? V : V2);  // MAXN/2 entries >= subgroups of size >= 2
    // head position of every slot: per-thread chunk, then an exclusive max-scan
    const uint32_t per = ceil_div(g.size, T);
    const uint32_t b0 = tid * per;
    uint32_t run = 0;
    for (uint32_t q = 0; q < per; ++q) {
        const uint32_t i = b0 + q;
        if (i < g.size) {
            if (i == 0 || Kf[i] != Kf[i - 1]) run = i;
            H[i] = run;
        }
    }
    const uint32_t excl = block_excl_max<WAVES>(run, wsum);
    for (uint32_t q = 0; q < per; ++q) {
        const uint32_t i = b0 + q;
        if (i < g.size && H[i] < excl) H[i] = excl;
    }
    __syncthreads();
    const uint32_t bo = cb ? bm.off[bm.find(g.start)] : 0u;
    for (uint32_t i = tid; i < g.size; i += T) {
        const uint32_t hp = H[i];
        const uint32_t v = Vf[i];
        sa[g.start + i] = v;
        if (hp) rank[sa_pos(v, cb, bo)] = g.start + hp;  // hp == 0: still the group's start
        const bool end = (i + 1 == g.size) || Kf[i + 1] != Kf[i];
        if (end && i > hp) {
            stage_push_e(S, E, g.start + hp, i - hp + 1, MAXN / 2);
            if (dep_out) dep_out[g.start + hp] = sub_depth(dep_in, g.start, Kf[i]);
        }
    }
    stage_flush_e(S, E, gl, MAXN / 2);
}

// ---- large groups (> 4096): block-segmented radix sort of key2, one -----
// segment per group. The groups of a round are taken in slot order and
// sorted where they lie: keys at their own SA slots in one buffer, the SA
// entries in sa, the other side of the ping-pong slot-indexed as well (tiles
// never span two groups; td.block is the group's index in lg, td.start a
// slot).
// mm[2g], mm[2g + 1]: the smallest and largest key of group g (cleared to
// all ones and zero before the launch). A group whose keys are all equal
// does not split this round: its members' successors all lie in one group.
__global__ __launch_bounds__(kThreads) void k_fill_large(const TileDesc* __restrict__ tiles,
                                                         const Grp* __restrict__ lg, const uint32_t* __restrict__ sa,
                                                         const uint32_t* __restrict__ rank, BlockMap bm, uint32_t cb,
                                                         uint32_t h, const uint32_t* __restrict__ dep,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ mm)
{
    __shared__ uint32_t wmin[kThreads / QZ_WAVE], wmax[kThreads / QZ_WAVE];
    const TileDesc td = tiles[blockIdx.x];
    const uint32_t g0 = lg[td.block].start;
    // a group lies in one block, whose text positions and SA slots are the
    // same range: keys relative to its start need only bits(block size)
    const uint32_t blk = bm.find(g0);
    const uint32_t bstart = bm.off[blk], bend = bm.off[blk + 1];
    const uint32_t d = dep ? dep[g0] : h;
    uint32_t kmin = 0xffffffffu, kmax = 0u;
    for (uint32_t i = threadIdx.x; i < td.len; i += kThreads) {
        const uint32_t slot = td.start + i;
        const uint64_t p = (uint64_t)sa_pos(sa[slot], cb, bstart) + d;
        const uint32_t k = (p < bend) ? rank[(uint32_t)p] - bstart + 1u : 0u;
        keys[slot] = k;
        kmin = min(kmin, k);
        kmax = max(kmax, k);
    }
    for (int o = 1; o < QZ_WAVE; o <<= 1) {
        kmin = min(kmin, (uint32_t)__shfl_xor(kmin, o, QZ_WAVE));
        kmax = max(kmax, (uint32_t)__shfl_xor(kmax, o, QZ_WAVE));
    }
    const uint32_t w = threadIdx.x / QZ_WAVE;
    if (lane_id() == 0) {
        wmin[w] = kmin;
        wmax[w] = kmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t v = 1; v < kThreads / QZ_WAVE; ++v) {
            kmin = min(kmin, wmin[v]);
            kmax = max(kmax, wmax[v]);
        }
        atomicMin(&mm[2 * td.block], kmin);
        atomicMax(&mm[2 * td.block + 1], kmax);
    }
}

__global__ __launch_bounds__(kThreads) void k_fill_minmax_init(uint32_t* __restrict__ mm, uint32_t n)
{
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < n) {
        mm[2 * i] = 0xffffffffu;
        mm[2 * i + 1] = 0u;
    }
}

// The large groups that do not split this round (mm from k_fill_large) go
// to the next round as they are: same members, same SA order and ranks, a
// depth grown by that of the group their keys point into.
__global__ __launch_bounds__(kThreads) void k_keep_large(const Grp* __restrict__ lg, uint32_t nlarge,
                                                         const uint32_t* __restrict__ mm, GroupLists gl, BlockMap bm,
                                                         const uint32_t* __restrict__ dep_in,
                                                         uint32_t* __restrict__ dep_out)
{
    __shared__ StageT<kThreads> S;
    stage_init(S);
    __syncthreads();
    const uint32_t i = blockIdx.x * kThreads + threadIdx.x;
    if (i < nlarge && mm[2 * i] == mm[2 * i + 1]) {
        const Grp g = lg[i];
        stage_push_t(S, g.start, g.size, kThreads);
        if (dep_out) {
            const uint32_t k = mm[2 * i], bstart = bm.off[bm.find(g.start)];
            dep_out[g.start] = sub_depth(dep_in, g.start, k ? k - 1u + bstart + 1u : 0u);
        }
    }
    stage_flush_t(S, gl, kThreads);
}

// Large-group finish, one radix tile per workgroup, over the sorted keys and
// entries at the groups' own slots. Same scheme as k_final0.
__global__ __launch_bounds__(kThreads) void k_final_large(const TileDesc* __restrict__ tiles,
                                                          const Grp* __restrict__ lg,
                                                          const uint32_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ vals,
                                                          const uint32_t* __restrict__ pre, uint32_t* __restrict__ sa,
                                                          uint32_t* __restrict__ rank, GroupLists gl, BlockMap bm,
                                                          const uint32_t* __restrict__ dep_in,
                                                          uint32_t* __restrict__ dep_out, uint32_t cb)
{
    constexpr uint32_t Q = kSortTile / (kThreads / QZ_WAVE);
    __shared__ Stage S;
    __shared__ uint32_t qlast[kThreads / QZ_WAVE];
    stage_init(S);
    const TileDesc td = tiles[blockIdx.x];
    const uint32_t g0 = lg[td.block].start;
    const uint32_t bstart = bm.off[bm.find(g0)];  // keys are block-local: rank + 1 - bstart
    const uint32_t lane = lane_id(), w = threadIdx.x / QZ_WAVE;
    const uint32_t qb = w * Q;
    uint32_t m = 0;
    for (uint32_t i = qb + lane; i < qb + Q && i < td.len; i += QZ_WAVE)
        if (slot_head(keys, td, i)) m = td.start + i;
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_xor(m, d, QZ_WAVE);
        m = o > m ? o : m;
    }
    if (lane == 0) qlast[w] = m;
    __syncthreads();
    uint32_t carry = pre[blockIdx.x];
    for (uint32_t v = 0; v < w; ++v) carry = qlast[v] > carry ? qlast[v] : carry;
    const bool last_tile = td.tlocal + 1 == td.ntiles;
    for (uint32_t b = qb; b < qb + Q && b < td.len; b += QZ_WAVE) {  // wave-uniform
        const uint32_t i = b + lane;
        const bool valid = i < td.len;
        const uint32_t j = td.start + i;
        const uint32_t k = valid ? keys[j] : 0u;
        const bool head = valid && slot_head(keys, td, i);
        uint32_t g = head ? j : 0u;
        for (int d = 1; d < QZ_WAVE; d <<= 1) {
            const uint32_t o = __shfl_up(g, d, QZ_WAVE);
            if (lane >= (uint32_t)d && o > g) g = o;
        }
        if (carry > g) g = carry;
        carry = __shfl(g, QZ_WAVE - 1, QZ_WAVE);
        if (valid) {
            const uint32_t s = vals[j];
            sa[j] = s;
            if (g != g0) rank[sa_pos(s, cb, bstart)] = g;  // the first subgroup keeps the group's start
            const bool end = (last_tile && i + 1 == td.len) || keys[j + 1] != k;
            if (end && j > g) {
                stage_push(S, g, j - g + 1);
                if (dep_out) dep_out[g] = sub_depth(dep_in, g0, k ? k - 1u + bstart + 1u : 0u);
            }
        }
    }
    stage_flush(S, gl);
}

// libsais-convention BWT of every block: U[0] = T[n-1]; U[j] = T[SA[j-1]-1]
// for 0 < j < p, T[SA[j]-1] for j >= p, with p = 1 + rank of suffix 0.
// The text as 4-bit symbol codes, eight per word, when the batch has at most
// 15 distinct bytes (sigma + 1 <= 16): k_bwt_out's random gathers then read
// half the bytes, so twice as much of a block's text stays in L2. A larger
// alphabet leaves the kernel idle and k_bwt_out reads the bytes.
// The 4-bit text: nib[i / 8] holds the codes of positions [i, i + 8), the
// first in the low nibble. for_fuse: built for the fused first doubling
// round's keys (nib_keys), else for the BWT output gathers (not needed in
// char mode).

__global__ __launch_bounds__(kThreads) void k_pack_nib(const uint8_t* __restrict__ text, uint32_t n,
                                                       const PackParams* __restrict__ ppd, uint32_t* __restrict__ nib,
                                                       uint32_t for_fuse)
{
    // uniform over the grid
    if (for_fuse ? !nib_keys(*ppd) : (ppd->radix > 16u || ppd->cb)) return;
    __shared__ uint8_t code[256];
    code[threadIdx.x] = (uint8_t)ppd->code[threadIdx.x];
    __syncthreads();
    const uint64_t i0 = ((uint64_t)blockIdx.x * kThreads + threadIdx.x) * 8u;  // text bytes [i0, i0 + 8)
    if (i0 >= n) return;
    uint32_t w = 0;
    if (i0 + 8u <= n && ((uintptr_t)(text + i0) & 7u) == 0) {
        const uint2 v = *reinterpret_cast<const uint2*>(text + i0);
#pragma unroll
        for (uint32_t b = 0; b < 4; ++b) {
            w |= (uint32_t)code[(v.x >> (8 * b)) & 255u] << (4 * b);
            w |= (uint32_t)code[(v.y >> (8 * b)) & 255u] << (16 + 4 * b);
        }
    } else {
        for (uint32_t b = 0; b < 8u; ++b)
            if (i0 + b < n) w |= (uint32_t)code[text[i0 + b]] << (4 * b);
    }
    nib[i0 / 8u] = w;
}

__global__ __launch_bounds__(kThreads) void k_bwt_out(const uint8_t* __restrict__ text,
                                                      const uint32_t* __restrict__ nib,
                                                      const PackParams* __restrict__ ppd,
                                                      const uint32_t* __restrict__ sa,
                                                      const uint32_t* __restrict__ rank,
                                                      BlockMap bm,
                                                      uint32_t n, uint8_t* __restrict__ out)
{
    if (ppd->cb) return;  // uniform: char mode is k_bwt_out_chars
    // code -> byte for the packed form (k_pack_nib ran with the same test)
    __shared__ uint8_t dec[16];
    const bool packed = ppd->radix <= 16u;
    if (packed) {
        const uint32_t cd = ppd->code[threadIdx.x];
        if (cd != 0u && cd < 16u) dec[cd] = (uint8_t)threadIdx.x;
    }
    __syncthreads();
    const uint32_t j = blockIdx.x * kThreads + threadIdx.x;
    if (j >= n) return;
    const uint32_t b = bm.find(j);
    const uint32_t o = bm.off[b], e = bm.off[b + 1];
    const uint32_t local = j - o;
    const uint32_t p = rank[o] - o + 1u;
    const uint32_t x = local == 0 ? e - 1u : (local < p ? sa[j - 1] : sa[j]) - 1u;
    out[j] = packed ? dec[(nib[x >> 3] >> (4u * (x & 7u))) & 15u] : text[x];
}

// The BWT in char mode: every SA entry holds the code of the symbol before
// its suffix, so the output is one streaming pass, 16 slots per thread (16-byte
// loads of entries, one 16-byte store when the output is aligned).
constexpr uint32_t kOutItems = 16;
__global__ __launch_bounds__(kThreads) void k_bwt_out_chars(const uint8_t* __restrict__ text,
                                                            const PackParams* __restrict__ ppd,
                                                            const uint32_t* __restrict__ sa,
                                                            const uint32_t* __restrict__ rank, BlockMap bm, uint32_t n,
                                                            uint8_t* __restrict__ out)
{
    if (!ppd->cb) return;  // uniform: k_bwt_out serves the batch
    __shared__ uint8_t dec[64];
    const uint32_t cmask = (1u << ppd->cb) - 1u;
    {
        const uint32_t cd = ppd->code[threadIdx.x];
        if (cd != 0u && cd <= cmask) dec[cd] = (uint8_t)threadIdx.x;
    }
    __syncthreads();
    const uint32_t j0 = (blockIdx.x * kThreads + threadIdx.x) * kOutItems;
    if (j0 >= n) return;
    uint32_t b = bm.find(j0);
    uint32_t o = bm.off[b], e = bm.off[b + 1];
    uint32_t p = rank[o] - o + 1u;  // the block's primary index
    uint32_t prev = j0 > o ? sa[j0 - 1] : 0u;
    uint32_t ent[kOutItems];
    if (j0 + kOutItems <= n) {
        const uint4* v = reinterpret_cast<const uint4*>(sa + j0);
#pragma unroll
        for (uint32_t k = 0; k < kOutItems / 4; ++k) {
            const uint4 x = v[k];
            ent[4 * k] = x.x;
            ent[4 * k + 1] = x.y;
            ent[4 * k + 2] = x.z;
            ent[4 * k + 3] = x.w;
        }
    } else {
#pragma unroll
        for (uint32_t q = 0; q < kOutItems; ++q) ent[q] = j0 + q < n ? sa[j0 + q] : 0u;
    }
    uint32_t w[kOutItems / 4] = {0, 0, 0, 0};
#pragma unroll
    for (uint32_t q = 0; q < kOutItems; ++q) {
        const uint32_t j = j0 + q;
        uint32_t c = 0;
        if (j < n) {
            if (j >= e) {  // a new block starts inside this thread's slots
                ++b;
                o = e;
                e = bm.off[b + 1];
                p = rank[o] - o + 1u;
            }
            const uint32_t local = j - o;
            c = local == 0 ? text[e - 1u] : dec[(local < p ? prev : ent[q]) & cmask];
            prev = ent[q];
        }
        w[q / 4] |= c << (8 * (q % 4));
    }
    uint8_t* dst = out + j0;
    if (j0 + kOutItems <= n && ((uintptr_t)dst & 15u) == 0) {
        *reinterpret_cast<uint4*>(dst) = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
        for (uint32_t q = 0; q < kOutItems && j0 + q < n; ++q) dst[q] = (uint8_t)(w[q / 4] >> (8 * (q % 4)));
    }
}

// Per block: primary index and sampled indexes I[k] = 1 + rank(k*r).
__global__ void k_bwt_index(const uint32_t* __restrict__ rank, BlockMap bm,
                            int32_t* __restrict__ idx /* nblk x 256 */)
{
    const uint32_t b = blockIdx.x;
    if (b >= bm.nblk) return;
    const uint32_t o = bm.off[b], n = bm.off[b + 1] - o;
    uint32_t r = n / 8;
    r |= r >> 1; r |= r >> 2; r |= r >> 4; r |= r >> 8; r |= r >> 16;
    r = (r >> 1) + 1;
    const uint32_t cnt = (n - 1) / r;
    for (uint32_t k = threadIdx.x; k <= cnt && k < 256; k += blockDim.x)
        idx[(size_t)b * 256 + k] = (int32_t)(rank[o + k * r] - o + 1u);
}

// ---- alphabet classes ------------------------------------------------------
// Round 0 packs one alphabet for the whole batch, so a batch that mixes
// reads (5 symbols) with names (about 40) packs the reads' keys in radix
// 41: 5 symbols per key instead of 12, and no 4-bit text or char-mode
// entries. Keys are only ever compared within a block, so blocks of
// different classes can be sorted as separate sub-batches with their own
// packing. k_block_class samples each block (16 bytes every kClassSample)
// for the symbols present and the run boundaries that decide the key mode;
// the host groups blocks by class from that. The sample decides speed only:
// each sub-batch still counts its exact alphabet.
constexpr uint32_t kClassSample = 512;
constexpr uint32_t kClassWords = 10;  // per block: 8 presence words, run boundaries, positions compared
__global__ __launch_bounds__(kThreads) void k_block_class(const uint8_t* __restrict__ text, uint32_t n, BlockMap bm,
                                                          uint32_t* __restrict__ cls)
{
    const uint64_t i64 = ((uint64_t)blockIdx.x * kThreads + threadIdx.x) * kClassSample;
    const bool live = i64 < n;  // batches reach 4 GB: no 32-bit wrap
    const uint32_t i0 = live ? (uint32_t)i64 : 0u;
    const uint32_t b = live ? bm.find(i0) : 0xffffffffu;
    uint32_t m[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t cnt = 0, seen = 0;
    if (live) {
        // one 16-byte load: the aligned window that holds position i0 (it
        // holds a byte of the text, so it lies in the text's pages), the
        // positions of the window inside i0's block and the batch
        const uint32_t bs = bm.off[b], be = bm.off[b + 1];
        const uint32_t mis = (uint32_t)((uintptr_t)(text + i0) & 15u);
        const uint4 v = *reinterpret_cast<const uint4*>(text + i0 - mis);
        const uint32_t w4[4] = {v.x, v.y, v.z, v.w};
        const int64_t p0 = (int64_t)i0 - mis;
        uint32_t prev = 0;
        bool have = false;
#pragma unroll
        for (uint32_t q = 0; q < 16; ++q) {
            const int64_t p = p0 + q;
            if (p < (int64_t)bs || p >= (int64_t)be) continue;
            const uint32_t c = (w4[q >> 2] >> (8u * (q & 3u))) & 255u;
#pragma unroll
            for (uint32_t w = 0; w < 8; ++w) m[w] |= (c >> 5) == w ? 1u << (c & 31u) : 0u;
            if (have) {
                cnt += c != prev ? 1u : 0u;
                ++seen;
            }
            prev = c;
            have = true;
        }
    }
    const uint32_t b0 = __shfl(b, 0, QZ_WAVE);
    if (!__ballot(b != b0)) {  // the whole wave samples one block
        for (int d = 1; d < QZ_WAVE; d <<= 1) {
#pragma unroll
            for (uint32_t w = 0; w < 8; ++w) m[w] |= __shfl_xor(m[w], d, QZ_WAVE);
            cnt += __shfl_xor(cnt, d, QZ_WAVE);
            seen += __shfl_xor(seen, d, QZ_WAVE);
        }
        if (lane_id() != 0 || b0 == 0xffffffffu) return;
    } else if (!live) {
        return;
    }
    uint32_t* c = cls + (size_t)b * kClassWords;
#pragma unroll
    for (uint32_t w = 0; w < 8; ++w)
        if (m[w]) atomicOr(&c[w], m[w]);
    if (cnt) atomicAdd(&c[8], cnt);
    if (seen) atomicAdd(&c[9], seen);
}

// Copies byte runs between the batch and a class-ordered staging buffer:
// one workgroup per run of at most kCopyChunk bytes, any alignment. Stores
// are aligned 32-bit words; each word's source bytes come from the two
// aligned words that hold them (the second one always holds a byte of the
// run, so no load leaves the source's pages).
struct CopyRun {
    uint32_t dst, src, len;
};
constexpr uint32_t kCopyChunk = 65536;
__global__ __launch_bounds__(kThreads) void k_copy_runs(const CopyRun* __restrict__ runs,
                                                        const uint8_t* __restrict__ src, uint8_t* __restrict__ dst)
{
    const CopyRun r = runs[blockIdx.x];
    const uint8_t* s = src + r.src;
    uint8_t* d = dst + r.dst;
    const uint32_t head = min((uint32_t)((4u - ((uintptr_t)d & 3u)) & 3u), r.len);
    if (threadIdx.x < head) d[threadIdx.x] = s[threadIdx.x];
    const uint32_t nw = (r.len - head) / 4u;
    const uint8_t* s1 = s + head;
    uint32_t* d1 = reinterpret_cast<uint32_t*>(d + head);
    const uint32_t sh = (uint32_t)((uintptr_t)s1 & 3u) * 8u;
    const uint32_t* sw = reinterpret_cast<const uint32_t*>((uintptr_t)s1 & ~(uintptr_t)3);
    if (sh == 0) {
        for (uint32_t w = threadIdx.x; w < nw; w += kThreads) d1[w] = sw[w];
    } else {
        for (uint32_t w = threadIdx.x; w < nw; w += kThreads) d1[w] = (sw[w] >> sh) | (sw[w + 1] << (32u - sh));
    }
    for (uint32_t i = head + nw * 4u + threadIdx.x; i < r.len; i += kThreads) d[i] = s[i];
}

// ---- launch-shape autotuner ---------------------------------------------
// Knobs that change speed and never a result, each chosen per process by
// timing its candidates on the first calls with enough work:
//   0 round-0 digit width {8, 11}: 11 bits saves a pass on 32-bit keys,
//     8 bits keeps digit runs long enough for coalesced stores;
//   1 large-group digit width {8, 9}: block-local keys need at most 27 bits
//     for blocks of up to 128 MiB (3 passes where 8-bit digits take 4; the
//     histograms have a buffer sized for 9-bit digits);
//   2 tile shape of the fused round-0 finish {512, 1024 threads per 4096-slot
//     tile}, both with the tile's suffixes held in LDS (the shapes without
//     them, and 256 threads, measured no better);
//   3 LDS budget of k_sort_small {16 KB, none}: extra LDS per workgroup only
//     caps how many of its random-access waves share a CU.
// A knob rotates its candidates call by call, timed with HIP events on the
// engine's stream around its own kernels only, and keeps the fastest per
// unit of work once every candidate has run. QZBSC_RADIX_BITS=8|9|11 (both
// widths; at most 9 for large groups), QZBSC_FUSE_SHAPE and
// QZBSC_LDS_SORT_SMALL force a choice.
static int fuse_shape();
enum TuneKnob { kKnobRound0Bits, kKnobLargeBits, kKnobFuseShape, kKnobSortSmallLds, kKnobCount };
constexpr int kLargeRadixBits = 9;  // the large-group sort's widest digit (its histograms have their own buffer)
constexpr int kKnobMaxCand = 4;
struct Knob {
    std::mutex mu;
    int choice = -1;  // candidate index, -1 while undecided
    double ns[kKnobMaxCand] = {0, 0, 0, 0};
    double units[kKnobMaxCand] = {0, 0, 0, 0};
    int trials = 0;
    int ncand = 0;
    int cand[kKnobMaxCand] = {0, 0, 0, 0};
    uint64_t min_units = 0;  // calls with less work are neither timed nor rotated
};
Knob g_knob[kKnobCount];
static std::once_flag g_tune_once;

static void tune_init()
{
    std::call_once(g_tune_once, [] {
        auto set = [](int k, std::initializer_list<int> c, uint64_t mn) {
            g_knob[k].ncand = 0;
            for (int v : c) g_knob[k].cand[g_knob[k].ncand++] = v;
            g_knob[k].min_units = mn;
        };
        set(kKnobRound0Bits, {8, 11}, 1u << 22);      // elements
        set(kKnobLargeBits, {8, 9}, 1u << 22);        // elements
        set(kKnobFuseShape, {0, 5}, 1u << 22);        // fuse_shape() codes; elements
        set(kKnobSortSmallLds, {16384, 0}, 1u << 16);  // waves of small-class work per call
    });
}

// The forced value of a knob (from its environment variable), or -1.
static int tune_forced(int which)
{
    const char* e = nullptr;
    switch (which) {
        case kKnobRound0Bits:
        case kKnobLargeBits:
            if ((e = getenv("QZBSC_RADIX_BITS"))) {
                const int v = atoi(e);
                if (v == 8 || v == 9 || v == 11) return v;
            }
            return -1;
        case kKnobFuseShape: return getenv("QZBSC_FUSE_SHAPE") ? fuse_shape() : -1;
        case kKnobSortSmallLds: return getenv(kOccEnv[kOccSortSmall]) ? (int)occ_lds(kOccSortSmall) : -1;
        default: return -1;
    }
}

// The value to use for a call with `units` of work.
static int tune_pick(int which, uint64_t units)
{
    tune_init();
    const int f = tune_forced(which);
    if (f >= 0) return f;
    Knob& t = g_knob[which];
    std::lock_guard<std::mutex> lk(t.mu);
    if (t.choice >= 0) return t.cand[t.choice];
    if (units < t.min_units) return t.cand[0];
    return t.cand[t.trials % t.ncand];
}

// True while the knob still times calls with `units` of work.
static bool tune_wants(int which, uint64_t units)
{
    tune_init();
    if (tune_forced(which) >= 0) return false;
    Knob& t = g_knob[which];
    std::lock_guard<std::mutex> lk(t.mu);
    return t.choice < 0 && units >= t.min_units;
}

// One timed call with candidate `value` over `units` of work; after a trial
// of every candidate the fastest per unit is kept for the process.
static void tune_report(int which, int value, uint64_t units, float ms)
{
    Knob& t = g_knob[which];
    std::lock_guard<std::mutex> lk(t.mu);
    if (t.choice >= 0 || units < t.min_units || ms <= 0) return;
    int k = -1;
    for (int c = 0; c < t.ncand; ++c)
        if (t.cand[c] == value) k = c;
    if (k < 0) return;
    t.ns[k] += ms * 1e6;
    t.units[k] += (double)units;
    ++t.trials;
    int best = -1;
    double best_ns = 0;
    for (int c = 0; c < t.ncand; ++c) {
        if (t.units[c] <= 0) return;  // a candidate is still untried
        const double v = t.ns[c] / t.units[c];
        if (best < 0 || v < best_ns) {
            best = c;
            best_ns = v;
        }
    }
    t.choice = best;
}

int tune_knob(int which, int* choice, int* ncand, int* cand, double* ns_per_unit)
{
    tune_init();
    Knob& t = g_knob[which];
    std::lock_guard<std::mutex> lk(t.mu);
    *choice = t.choice >= 0 ? t.cand[t.choice] : -1;
    *ncand = t.ncand;
    for (int c = 0; c < t.ncand; ++c) {
        cand[c] = t.cand[c];
        ns_per_unit[c] = t.units[c] > 0 ? t.ns[c] / t.units[c] : 0;
    }
    return t.trials;
}

int tune_state(int which, int* choice, double* ns_per_elem)
{
    tune_init();
    Knob& t = g_knob[which];
    std::lock_guard<std::mutex> lk(t.mu);
    *choice = t.choice >= 0 ? t.cand[t.choice] : 0;
    for (int k = 0; k < 3; ++k) ns_per_elem[k] = 0;
    for (int c = 0; c < t.ncand; ++c) {
        const int w = t.cand[c] == 9 ? 1 : t.cand[c] == 11 ? 2 : 0;
        ns_per_elem[w] = t.units[c] > 0 ? t.ns[c] / t.units[c] : 0;
    }
    return t.trials;
}

// The SA entries of round 0's first radix pass, made from the text instead
// of read (sa_entry of each position; tiles never span two blocks). The
// code table and the char-mode flag are decided on the device, so they are
// read from the packing parameters into LDS.
struct SaEntrySource {
    const uint8_t* text;
    const PackParams* pp;
    const uint32_t* off;  // block offsets
    struct Lds {
        uint16_t code[256];
        uint32_t cb;
    };
    __device__ __forceinline__ void stage(Lds& l) const
    {
        for (uint32_t d = threadIdx.x; d < 256u; d += blockDim.x) l.code[d] = pp->code[d];
        if (threadIdx.x == 0) l.cb = pp->cb;
    }
    template <int IT>
    __device__ __forceinline__ void load(const Lds& l, const TileDesc& td, uint32_t e0, uint32_t last,
                                         uint32_t (&vv)[IT]) const
    {
        const uint32_t bo = off[td.block];
        uint8_t prev[IT];
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const uint32_t i = min(td.start + e0 + (uint32_t)it * QZ_WAVE, last);
            prev[it] = text[i > bo ? i - 1u : i];
        }
#pragma unroll
        for (int it = 0; it < IT; ++it) {
            const uint32_t i = min(td.start + e0 + (uint32_t)it * QZ_WAVE, last);
            vv[it] = sa_entry(i, bo, l.cb, l.code, prev[it]);
        }
    }
};

template <typename K, class VS0>
static int seg_sort_bits(int bits, K* k0, uint32_t* v0, K* k1, uint32_t* v1, const TileDesc* tiles, uint32_t nt,
                         uint32_t hist_words, int key_bits, uint32_t* hist, uint32_t* stmp, hipStream_t s, bool* alt,
                         VS0 first, const K* kin = nullptr, const int32_t* seg_delta = nullptr)
{
    if (bits == 11)
        return seg_radix_sort<K, 11, VS0>(k0, v0, k1, v1, tiles, nt, hist_words, key_bits, hist, stmp, s, alt, first,
                                          kin, seg_delta);
    if (bits == 9)
        return seg_radix_sort<K, 9, VS0>(k0, v0, k1, v1, tiles, nt, hist_words, key_bits, hist, stmp, s, alt, first,
                                         kin, seg_delta);
    return seg_radix_sort<K, 8, VS0>(k0, v0, k1, v1, tiles, nt, hist_words, key_bits, hist, stmp, s, alt, first, kin,
                                     seg_delta);
}

// Radix passes of a key_bits-bit sort with bits-bit digits.
static int radix_passes(int key_bits, int bits) { return (key_bits + bits - 1) / bits; }

static int bitlen(uint64_t v)
{
    int b = 0;
    while (v) { ++b; v >>= 1; }
    return b;
}

// Counter slots in cnt_ (slot i at word i * kCtrStride): set A classes
// [0, 9) and pool top 9, set B classes [32, 41) and top 41, overflow flag 18,
// bucket fill counters [48, 57).
constexpr int kCntB = 32, kTopOff = kClasses, kOverflow = 2 * kClasses, kFill = 48;
constexpr int kCtrSlots = kFill + kClasses;
constexpr size_t kCtrWords = (size_t)kCtrSlots * kCtrStride;
constexpr uint32_t kPinnedPack = 4096;  // pinned_ word of the PackParams read back after round 0

static uint64_t pool_cap(uint64_t n) { return n / 2 + 1; }
// Copy runs of one direction of a class split: blocks cut at kCopyChunk.
static uint64_t copy_runs_cap(uint64_t n, uint32_t nblk) { return n / kCopyChunk + nblk + 1; }

// The sort workspace covers at most kChunkDefault bytes of whole blocks (or
// the largest block, if larger); a bigger batch is sorted chunk by chunk in
// it, and only the text and output staging scale with the batch. Chunks of
// 512 MiB cost 0-3% of forward time against one batch (1.5 GB of reads,
// binned qualities, a mixed batch), chunks of 256 MiB 2-8%.
// QZBSC_CHUNK_MIB sets the chunk (0: none).
constexpr uint64_t kChunkDefault = 512ull << 20;
static uint64_t chunk_bytes()
{
    const char* e = getenv("QZBSC_CHUNK_MIB");
    return e ? (uint64_t)std::max(0L, atol(e)) << 20 : kChunkDefault;
}
uint64_t BwtEngine::chunk() { return chunk_bytes(); }

// Workspace elements for a batch of n bytes whose largest block has maxblk.
// A chunk stays below 2^31 bytes (positions and SA entries are u32, a
// batch of the BWT entries may hold up to 4 GB).
constexpr uint64_t kChunkMax = 0x7fff0000ull;
static uint64_t work_cap(uint64_t n, uint64_t maxblk)
{
    const uint64_t c = chunk_bytes();
    const uint64_t w = c == 0 ? n : std::min<uint64_t>(n, std::max<uint64_t>(c, maxblk));
    return std::min<uint64_t>(w, std::max<uint64_t>(kChunkMax, maxblk));
}

uint64_t BwtEngine::footprint(uint64_t n, uint32_t nblk, uint64_t maxblk)
{
    const uint64_t c = std::max<uint64_t>(n, 1024);
    const uint32_t cb = std::max<uint32_t>(nblk, 1);
    const uint64_t w = std::max<uint64_t>(work_cap(c, maxblk), 1024);
    const size_t seg_words = ((size_t)max_tiles(w, cb) + w / kSortTile + 2) * (1u << kMaxRadixBits);
    const size_t tmpw = scan_tmp_words(seg_words) + 1024;
    const uint64_t groups = pool_cap(w);
    const size_t tile_cap = (size_t)max_tiles(w, cb) + w / kSortTile + 2;
    const uint64_t nlg = w / (kMedMax + 1) + 2;  // large groups of a round, at most
    // staging by the batch: text, out, block table, copy runs, per-block words;
    // workspace by the chunk: sa, rank, keys0, vals0 (4 B each); dep0/1 (4 B
    // each, allocated for run-length batches only, counted always); bufB
    // (doubling keys or round-0 histograms); bufA per tile; large-group
    // histograms (9-bit digits at most), their list, offsets, shifts, key ranges
    return 2 * (c + 64) + ((c >> kTabShift) + 2) * 4 + 2 * copy_runs_cap(c, cb) * sizeof(CopyRun) +
           3 * (cb + 1) * 4 + (uint64_t)cb * 1024 + (uint64_t)cb * kClassWords * 4 + 1024 + 256 +
           sizeof(PackParams) + 4 * 4 * w + 2 * 4 * w + std::max<uint64_t>(w, seg_words) * 4 + tile_cap * 4 +
           tmpw * 4 + 2 * groups * sizeof(Grp) + nlg * (sizeof(Grp) + 4 + 8) +
           tile_cap * (1u << kLargeRadixBits) * 4 + 2 * tile_cap * sizeof(TileDesc);
}

// Depth arrays, allocated for the engine's first batch with run-length keys
// and kept from then on.
int BwtEngine::ensure_depths()
{
    if (dep_[0]) return 0;
    QZ_HIP(hipMalloc(&dep_[0], wcap_ * 4));
    QZ_HIP(hipMalloc(&dep_[1], wcap_ * 4));
    return 0;
}

int BwtEngine::reserve(uint64_t n, uint32_t nblk, uint64_t maxblk)
{
    const uint64_t wn = work_cap(std::max<uint64_t>(n, 1024), std::min<uint64_t>(maxblk, n));
    if (n <= cap_ && nblk <= cap_blk_ && wn <= wcap_) return 0;
    const MemTrace mt("forward", n, nblk);
    // grow geometrically so a stream of slowly growing batches reallocates
    // only a few times
    const uint64_t c = std::max<uint64_t>(std::max<uint64_t>(n, 1024), cap_ ? cap_ + cap_ / 2 : 0);
    const uint32_t cb = std::max<uint32_t>(std::max<uint32_t>(nblk, 1), cap_blk_ ? cap_blk_ + cap_blk_ / 2 : 0);
    const uint64_t w = std::min<uint64_t>(c, std::max<uint64_t>(wn, wcap_ ? wcap_ + wcap_ / 2 : 0));
    release();
    const size_t seg_words = ((size_t)max_tiles(w, cb) + w / kSortTile + 2) * (1u << kMaxRadixBits);
    const size_t tmpw = scan_tmp_words(seg_words) + 1024;  // scan scratch; the histograms live in bufB
    const size_t tile_cap = (size_t)max_tiles(w, cb) + w / kSortTile + 2;  // round 0 or large groups
    QZ_HIP(hipMalloc(&text_, c + 64));
    QZ_HIP(hipMalloc(&out_, c + 64));
    QZ_HIP(hipMalloc(&sa_, w * 4));
    QZ_HIP(hipMalloc(&rank_, w * 4));
    QZ_HIP(hipMalloc(&bufA_, tile_cap * 4));  // last head per radix tile
    QZ_HIP(hipMalloc(&bufB_, std::max<size_t>(w, seg_words) * 4));
    QZ_HIP(hipMalloc(&keys0_, w * 4));
    QZ_HIP(hipMalloc(&vals0_, w * 4));
    QZ_HIP(hipMalloc(&lhist_, tile_cap * (1u << kLargeRadixBits) * 4));
    QZ_HIP(hipMalloc(&tmp_, tmpw * 4));
    QZ_HIP(hipMalloc(&off_, (cb + 1) * 4));
    QZ_HIP(hipMalloc(&idx_, (size_t)cb * 256 * 4));
    QZ_HIP(hipMalloc(&hist_, 256 * 4));
    QZ_HIP(hipMalloc(&brun_, 2 * ((size_t)cb + 1) * 4));
    QZ_HIP(hipMalloc(&cnt_, kCtrWords * 4));
    for (int set = 0; set < 2; ++set) QZ_HIP(hipMalloc(&groups_[set], pool_cap(w) * sizeof(Grp)));  // pool, list
    QZ_HIP(hipMalloc(&lgrp_, (w / (kMedMax + 1) + 2) * sizeof(Grp)));
    QZ_HIP(hipMalloc(&ldelta_, (w / (kMedMax + 1) + 2) * 4));
    QZ_HIP(hipMalloc(&lmm_, (w / (kMedMax + 1) + 2) * 8));
    QZ_HIP(hipHostMalloc((void**)&lmm_host_, (w / (kMedMax + 1) + 2) * 8));
    if (!lev_) QZ_HIP(hipEventCreateWithFlags(&lev_, hipEventDisableTiming));
    QZ_HIP(hipMalloc(&tiles_, tile_cap * sizeof(TileDesc)));
    QZ_HIP(hipMalloc(&tiles0_, tile_cap * sizeof(TileDesc)));
    QZ_HIP(hipMalloc(&ppd_, sizeof(PackParams)));
    QZ_HIP(hipMalloc(&btab_, ((c >> kTabShift) + 2) * 4));
    QZ_HIP(hipMalloc(&cls_, (size_t)cb * kClassWords * 4));
    QZ_HIP(hipMalloc(&runs_, 2 * copy_runs_cap(c, cb) * sizeof(CopyRun)));
    tiles_host_.resize(tile_cap);
    QZ_HIP(hipHostMalloc(&pinned_, (kPinnedPack + 4) * 4));  // counter copy, then the packing words
    for (auto& ev : tev_)
        if (!ev) QZ_HIP(hipEventCreate(&ev));
    cap_ = c;
    cap_blk_ = cb;
    wcap_ = w;
    return 0;
}

void BwtEngine::release()
{
    void* ps[] = {text_,      out_,       sa_,    rank_,  bufA_,   bufB_, keys0_, vals0_, dep_[0], dep_[1], tmp_,
                  off_,       idx_,       hist_,  brun_,  cnt_,    groups_[0], groups_[1], lgrp_,   ldelta_,
                  lhist_,     tiles_,     tiles0_, ppd_,  btab_,  cls_,    runs_,  lmm_};
    g0_.reset();
    for (void* p : ps)
        if (p) (void)hipFree(p);
    if (pinned_) (void)hipHostFree(pinned_);
    if (lmm_host_) (void)hipHostFree(lmm_host_);
    lmm_ = lmm_host_ = nullptr;
    text_ = out_ = nullptr;
    sa_ = rank_ = bufA_ = bufB_ = vals0_ = tmp_ = off_ = hist_ = brun_ = cnt_ = lhist_ = nullptr;
    keys0_ = nullptr;
    dep_[0] = dep_[1] = nullptr;
    lgrp_ = nullptr;
    ldelta_ = nullptr;
    groups_[0] = groups_[1] = nullptr;
    tiles_ = tiles0_ = ppd_ = nullptr;
    btab_ = nullptr;
    cls_ = nullptr;
    runs_ = nullptr;
    idx_ = nullptr;
    pinned_ = nullptr;
    cap_ = 0;
    cap_blk_ = 0;
    wcap_ = 0;
}

static GroupLists make_lists(void* pool, uint32_t* cnt, uint32_t* overflow, uint64_t cap, bool down)
{
    GroupLists gl;
    gl.pool = (Grp*)pool;
    gl.cnt = cnt;
    gl.top = cnt + kTopOff * kCtrStride;
    gl.cap = (uint32_t)pool_cap(cap);
    gl.overflow = overflow;
    gl.down = down ? 1u : 0u;
    return gl;
}

// QZBSC_RLE=0|1 forces packed or run-length round-0 keys (default: by the
// batch's mean run length).
static int rle_force()
{
    const char* e = getenv("QZBSC_RLE");
    return e ? (atoi(e) ? 1 : 0) : -1;
}

// Round 0 as one launch sequence with no host round trip, so it can be
// replayed as a graph: clear the counters, count symbols, derive the packing
// on the device, build the packed prefix keys, sort them (32 key bits),
// then mark group heads and emit SA, ranks and the unsorted groups.
// QZBSC_FUSE=0 keeps the first doubling round out of round 0's finish.
static bool fuse_enabled()
{
    const char* e = getenv("QZBSC_FUSE");
    return !(e && e[0] == '0');
}

// Launch shape of the fused finish, QZBSC_FUSE_SHAPE=<threads>[s]: 256, 512
// or 1024 threads per tile, "s" for suffixes kept in LDS. Returns the code
// enqueue_round0 takes (0 = the default, 512 threads with suffixes in LDS,
// the fastest on the MI355X).
static int fuse_shape()
{
    const char* e = getenv("QZBSC_FUSE_SHAPE");
    if (!e) return 0;
    const std::string v = e;
    return v == "256" ? 1 : v == "256s" ? 2 : v == "512" ? 3 : v == "1024" ? 4 : v == "1024s" ? 5 : 0;
}

int BwtEngine::enqueue_round0(const uint8_t* text, uint32_t n, uint32_t nblk, int bits, uint32_t ntiles,
                              uint32_t hist_words, bool timed, uint32_t fuse, int fshape, hipStream_t s)
{
    const BlockMap bmap_{off_, btab_, nblk};
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kThreads), 0, s, hist_, 256u);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kThreads), 0, s, brun_, 2 * nblk);
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(kThreads), 0, s, cnt_, (uint32_t)kCtrWords);
    QZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_byte_hist, dim3(std::min<uint32_t>(qz_div_up(n, kThreads), 2048)), dim3(kThreads), 0, s,
                       text, n, hist_);
    QZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_block_runs, dim3(qz_div_up(n, kThreads * kRunSample)), dim3(kThreads), 0, s, text, n, bmap_,
                       brun_);
    QZ_HIP(hipGetLastError());
    uint32_t maxb = 0;
    for (uint32_t b = 0; b < nblk; ++b) maxb = std::max(maxb, offs_host_[b + 1] - offs_host_[b]);
    const uint32_t maxb_bits = (uint32_t)bitlen(maxb);  // bits of the largest block-local position + 1
    hipLaunchKernelGGL(k_make_pack, dim3(1), dim3(kThreads), 0, s, hist_, (PackParams*)ppd_, rle_force(), bmap_,
                       (const uint32_t*)brun_, maxb_bits);
    QZ_HIP(hipGetLastError());
    // positional keys in the group list buffer (free until the first
    // doubling round): the radix sort reads them in its first pass and
    // leaves them intact for the fused first doubling round
    uint32_t* kpos = (uint32_t*)groups_[1];
    // the radix ping-pong: (keys0, vals0) and (group pool, sa), both free
    // until the finish writes sa, ranks and groups; the last pass lands in
    // (keys0, vals0), so with an odd pass count the first pass writes to it
    const bool odd = (radix_passes(32, bits) & 1) != 0;
    uint32_t* pk = (uint32_t*)groups_[0];
    hipLaunchKernelGGL(k_round0_keys, dim3(qz_div_up(n, kThreads * kKeyItems)), dim3(kThreads), 0, s, text, n, bmap_,
                        (const PackParams*)ppd_, (const uint32_t*)brun_, kpos, (uint32_t*)nullptr);
    QZ_HIP(hipGetLastError());
    bool alt = false;
    if (timed) QZ_HIP(hipEventRecord(tev_[0], s));
    // the first pass makes the SA entries from the text (k_round0_keys
    // writes only the keys)
    const SaEntrySource ents{text, (const PackParams*)ppd_, off_};
    if (odd)
        QZ_TRY(seg_sort_bits<uint32_t>(bits, pk, sa_, keys0_, vals0_, (const TileDesc*)tiles0_, ntiles, hist_words, 32,
                                       bufB_, tmp_, s, &alt, ents, kpos));
    else
        QZ_TRY(seg_sort_bits<uint32_t>(bits, keys0_, vals0_, pk, sa_, (const TileDesc*)tiles0_, ntiles, hist_words, 32,
                                       bufB_, tmp_, s, &alt, ents, kpos));
    if (timed) QZ_HIP(hipEventRecord(tev_[1], s));
    if (alt != odd) return QZ_GPU_ERROR;  // cannot happen: the last pass writes (keys0, vals0)
    uint32_t* keys = keys0_;
    uint32_t* vals = vals0_;
    // group starts: last head per tile, an exclusive max-scan over tiles,
    // then tile-local scans inside the finishing kernel
    hipLaunchKernelGGL(k_tile_lasthead, dim3(ntiles), dim3(kThreads), 0, s, keys, (const TileDesc*)tiles0_, bufA_);
    QZ_HIP(hipGetLastError());
    QZ_TRY((device_scan<OpMax, false>(bufA_, bufA_, ntiles, tmp_, s)));
    const GroupLists glA = make_lists(groups_[0], cnt_, cnt_ + kOverflow * kCtrStride, wcap_, false);
    const GroupLists glB = make_lists(groups_[0], cnt_ + kCntB * kCtrStride, cnt_ + kOverflow * kCtrStride, wcap_, true);
    // one of the two finishes the batch (the other returns at once): the
    // choice depends on the key mode and alphabet, decided on the device
    hipLaunchKernelGGL(k_final0, dim3(ntiles), dim3(kThreads), occ_lds(kOccFinal0), s, keys, vals,
                       (const TileDesc*)tiles0_, bufA_, sa_, rank_, glA, (const PackParams*)ppd_, bmap_, fuse);
    QZ_HIP(hipGetLastError());
    if (fuse) {
        // the 4-bit text for the fused round's keys, in bufB (the radix
        // histograms are done with); the kernel returns at once when the
        // keys cannot use it
        hipLaunchKernelGGL(k_pack_nib, dim3(qz_div_up(n, kThreads * 8u)), dim3(kThreads), 0, s, text, n,
                           (const PackParams*)ppd_, bufB_, 1u);
        QZ_HIP(hipGetLastError());
        if (timed) QZ_HIP(hipEventRecord(tev_[4], s));
#define QZ_FINAL0F(T, SVL)                                                                                      \
    hipLaunchKernelGGL((k_final0f<T, SVL>), dim3(ntiles), dim3(T), 0, s, keys, vals, (const TileDesc*)tiles0_, \
                       bufA_, sa_, rank_, glA, glB, (const PackParams*)ppd_, bmap_, (const uint32_t*)kpos,      \
                       (const uint32_t*)bufB_, fuse)
        switch (fshape) {
            case 1: QZ_FINAL0F(256, false); break;
            case 2: QZ_FINAL0F(256, true); break;
            case 3: QZ_FINAL0F(512, false); break;
            case 4: QZ_FINAL0F(1024, false); break;
            case 5: QZ_FINAL0F(1024, true); break;
            default: QZ_FINAL0F(512, true); break;
        }
#undef QZ_FINAL0F
        QZ_HIP(hipGetLastError());
        if (timed) QZ_HIP(hipEventRecord(tev_[5], s));
    }
    return 0;
}

// Sorts the suffixes of the batch in text (device); off_ holds the offsets.
int BwtEngine::sort_suffixes(const uint8_t* text, uint32_t n, uint32_t nblk, hipStream_t s)
{
    const BlockMap bmap_{off_, btab_, nblk};
    rounds_ = 0;
    trace_ = getenv("QZBSC_TRACE") != nullptr;

    if (offs_host_.size() != (size_t)nblk + 1) return QZ_BAD_PARAMETER;
    uint32_t* overflow = cnt_ + kOverflow * kCtrStride;
    // emission targets: the pool with counter set A or B, alternating
    const GroupLists gls[2] = {make_lists(groups_[0], cnt_, overflow, wcap_, false),
                               make_lists(groups_[0], cnt_ + kCntB * kCtrStride, overflow, wcap_, true)};
    Grp* const list = (Grp*)groups_[1];

    // ---- round 0: block-segmented radix sort on the packed prefix --------
    const int bits0 = tune_pick(kKnobRound0Bits, n);
    const int fshape = tune_pick(kKnobFuseShape, n);
    // a knob of round 0 is still timing: run eagerly with events
    const bool timed0 = tune_wants(kKnobRound0Bits, n) || tune_wants(kKnobFuseShape, n);
    uint32_t ntiles = 0;
    const uint32_t hist_words = count_tiles(offs_host_.data(), nblk, &ntiles, 1u << bits0);
    // the fused first doubling round needs room for the 4-bit text in bufB
    const uint32_t fuse = fuse_enabled() ? 1u : 0u;
    auto prepare0 = [&](hipStream_t st) -> int {
        uint32_t nt = 0;
        build_tiles(offs_host_.data(), nblk, tiles_host_.data(), &nt, 1u << bits0);
        QZ_HIP(hipMemcpyAsync(tiles0_, tiles_host_.data(), nt * sizeof(TileDesc), hipMemcpyHostToDevice, st));
        return 0;
    };
    auto launch0 = [&](hipStream_t st) -> int {
        return enqueue_round0(text, n, nblk, bits0, ntiles, hist_words, timed0, fuse, fshape, st);
    };
    if (timed0) {
        g0_.drop();
        QZ_TRY(prepare0(s));
        QZ_TRY(launch0(s));
    } else {
        std::vector<uint64_t> key = {n, nblk, (uint64_t)(uintptr_t)text, (uint64_t)bits0, cap_, cap_blk_, wcap_,
                                     (uint64_t)(rle_force() + 1), (uint64_t)fuse, (uint64_t)fshape};
        key.insert(key.end(), offs_host_.begin(), offs_host_.end());
        QZ_TRY(g0_.run(key, s, prepare0, launch0));
    }
    ++rounds_;

    // ---- doubling rounds ---------------------------------------------------
    uint32_t maxb = 0;
    for (uint32_t b = 0; b < nblk; ++b) maxb = std::max(maxb, offs_host_[b + 1] - offs_host_[b]);
    const int lkey_bits = bitlen(maxb);  // block-local large-group keys are <= block size
    uint32_t h = 0;  // the doubling depth of packed keys: k0, read back with the first counts
    bool rle = false;  // run-length keys: every group carries its own depth (dep_)
    uint32_t cb = 0;   // SA entries in char mode (4) or global positions (0)
    int cur = 0;
    bool large_timed = false;
    int large_bits = 8;
    uint64_t large_n = 0;
    // k_sort_small's LDS budget: chosen at the call's first round with
    // small-class work, timed over all of the call's rounds
    int ss_lds = -1;
    bool ss_timed = false, ss_round = false;
    uint64_t ss_slots = 0;
    double ss_ms = 0;
    uint32_t* hc = (uint32_t*)pinned_;
    // large-group lists of a round; their uploads are read before the next
    // round's synchronisation, so they live outside the loop
    std::vector<Grp> lg, lg_split;
    std::vector<uint32_t> lb;
    std::vector<int32_t> delta;
    for (;;) {
        // one copy brings back the current set's class counts and the
        // overflow word, which lies between the two sets in cnt_
        // (after round 0 both sets: the fused first round may have filled B)
        const int lo = cur ? kOverflow : 0, hi = (cur || rounds_ == 1) ? kCntB + kClasses : kOverflow + 1;
        QZ_HIP(hipMemcpyAsync(hc, cnt_ + (size_t)lo * kCtrStride, ((size_t)(hi - 1 - lo) * kCtrStride + 1) * 4,
                              hipMemcpyDeviceToHost, s));
        if (rounds_ == 1)  // k0, rle, cb
            QZ_HIP(hipMemcpyAsync(hc + kPinnedPack, &((const PackParams*)ppd_)->k0, 12, hipMemcpyDeviceToHost, s));
        QZ_HIP(hipStreamSynchronize(s));
        // counter slot i of the copy
        auto slot = [&](int i) { return ((volatile uint32_t*)hc)[(size_t)(i - lo) * kCtrStride]; };
        if (rounds_ == 1) {
            h = ((volatile uint32_t*)hc)[kPinnedPack];
            rle = ((volatile uint32_t*)hc)[kPinnedPack + 1] != 0;
            cb = ((volatile uint32_t*)hc)[kPinnedPack + 2];
            k0_ = h;
            cb_ = cb;
        }
        const int cbase = cur ? kCntB : 0;
        uint32_t counts[kClasses];
        uint64_t total = 0;
        for (int k = 0; k < kClasses; ++k) {
            counts[k] = slot(cbase + k);
            total += counts[k];
        }
        if (slot(kOverflow)) {
            fprintf(stderr, "qzbsc: group list overflow\n");
            return QZ_GPU_ERROR;
        }
        if (timed0 && rounds_ == 1) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, tev_[0], tev_[1]) == hipSuccess) tune_report(kKnobRound0Bits, bits0, n, ms);
            // the fused finish works only on packed keys (else it returns at once)
            if (fuse && !rle && hipEventElapsedTime(&ms, tev_[4], tev_[5]) == hipSuccess)
                tune_report(kKnobFuseShape, fshape, n, ms);
        }
        if (large_timed) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, tev_[2], tev_[3]) == hipSuccess)
                tune_report(kKnobLargeBits, large_bits, large_n, ms);
            large_timed = false;
        }
        if (ss_round) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, tev_[6], tev_[7]) == hipSuccess) ss_ms += ms;
            ss_round = false;
        }
        uint64_t fused = 0;  // groups round 0 already took through the first doubling round (set B)
        if (rounds_ == 1)
            for (int k = 0; k < kClasses; ++k) fused += slot(kCntB + k);
        if (trace_) {
            if (rle)
                fprintf(stderr, "qzbsc: round %d (per-group depths) groups", rounds_);
            else
                fprintf(stderr, "qzbsc: round %d h=%u groups", rounds_, h);
            for (int k = 0; k < kClasses; ++k) fprintf(stderr, " %u", counts[k]);
            if (fused) fprintf(stderr, " (+%llu for round 2)", (unsigned long long)fused);
            fprintf(stderr, "\n");
        }
        if (total == 0) {
            if (!fused) {
                if (ss_timed) tune_report(kKnobSortSmallLds, ss_lds, ss_slots, (float)ss_ms);
                break;
            }
            // every group round 0 left was fused: the first doubling round
            // has nothing to sort, and set B holds the second round's groups
            ++rounds_;
            h = (h > 0x40000000u) ? 0x80000000u : h * 2;
            cur ^= 1;
            continue;
        }
        if (rounds_ > 40) {
            fprintf(stderr, "qzbsc: suffix sort did not converge\n");
            return QZ_GPU_ERROR;
        }
        const GroupLists& C = gls[cur];
        const GroupLists& N = gls[cur ^ 1];
        if (rounds_ == 1 && rle) {
            // run-length keys: per-group depths, from round 0's groups (set A)
            QZ_TRY(ensure_depths());
            hipLaunchKernelGGL(k_init_depths, dim3(qz_div_up(total, kThreads)), dim3(kThreads), 0, s,
                               (const Grp*)C.pool, (uint32_t)total, keys0_, (const PackParams*)ppd_, bmap_,
                               (const uint32_t*)brun_, dep_[0]);
            QZ_HIP(hipGetLastError());
        }
        // the last round's groups: pool -> list ordered by class
        CurLists cl;
        {
            uint32_t o = 0;
            for (int k = 0; k < kClasses; ++k) {
                cl.list[k] = list + o;
                o += counts[k];
            }
            cl.cnt = C.cnt;
            QZ_HIP(hipMemsetAsync(cnt_ + kFill * kCtrStride, 0, ((kClasses - 1) * kCtrStride + 1) * 4, s));
            hipLaunchKernelGGL(k_bucket, dim3(qz_div_up(o, kBucketRange)), dim3(256), 0, s, C.pool, o, C.down, C.cap,
                               C.cnt, list, cnt_ + kFill * kCtrStride);
            QZ_HIP(hipGetLastError());
        }
        // this round emits into the pool again, with the other counter set
        // (after round 0 that set may already hold fused groups: append)
        if (rounds_ != 1) QZ_HIP(hipMemsetAsync(N.cnt, 0, (kClasses * kCtrStride + 1) * 4, s));  // classes and top

        SmallPlan plan;
        plan.wave_base[0] = 0;
        for (int k = 0; k < kSmallClasses; ++k) {
            const uint32_t per_wave = QZ_WAVE / (2u << k);
            plan.wave_base[k + 1] = plan.wave_base[k] + qz_div_up(counts[k], per_wave);
        }
        const uint32_t small_waves = plan.wave_base[kSmallClasses];
        const uint32_t small_sort_blocks = qz_div_up(small_waves, (kThreads / QZ_WAVE) * kSmallTasks);
        const uint32_t nwave = counts[kClsWave];
        const uint32_t nmed = counts[kClsMed];
        const uint32_t nlarge = counts[kClsLarge];

        // large groups: in slot order, host-side prefix of sizes, per-group
        // tiles at the groups' own slots
        uint32_t M = 0, lt = 0, lhist = 0;
        int lbits = 8;
        if (nlarge) {
            lg.resize(nlarge);
            QZ_HIP(hipMemcpyAsync(lg.data(), cl.list[kClsLarge], nlarge * sizeof(Grp), hipMemcpyDeviceToHost, s));
            QZ_HIP(hipStreamSynchronize(s));
            std::sort(lg.begin(), lg.end(), [](const Grp& a, const Grp& b) { return a.start < b.start; });
            lb.resize(nlarge + 1);
            for (uint32_t i = 0; i < nlarge; ++i) {
                lb[i] = M;
                M += lg[i].size;
            }
            lb[nlarge] = M;
            if (trace_) {
                uint64_t sz[5] = {0, 0, 0, 0, 0};  // elements in groups <= 8K, 16K, 64K, 1M, larger
                uint32_t mx = 0;
                for (uint32_t i = 0; i < nlarge; ++i) {
                    const uint32_t z = lg[i].size;
                    sz[z <= 8192 ? 0 : z <= 16384 ? 1 : z <= 65536 ? 2 : z <= (1u << 20) ? 3 : 4] += z;
                    mx = std::max(mx, z);
                }
                fprintf(stderr, "qzbsc:   large M=%u max=%u by size <=8K %llu <=16K %llu <=64K %llu <=1M %llu more %llu\n", M,
                        mx, (unsigned long long)sz[0], (unsigned long long)sz[1], (unsigned long long)sz[2],
                        (unsigned long long)sz[3], (unsigned long long)sz[4]);
            }
            lbits = std::min(tune_pick(kKnobLargeBits, M), kLargeRadixBits);
            lhist = build_tiles(lb.data(), nlarge, tiles_host_.data(), &lt, 1u << lbits);
            // tiles at the groups' slots; the sort writes each group shifted
            // from its packed offset to its slot start
            delta.resize(nlarge);
            for (uint32_t i = 0; i < nlarge; ++i) delta[i] = (int32_t)(lg[i].start - lb[i]);
            for (uint32_t t = 0; t < lt; ++t) tiles_host_[t].start += (uint32_t)delta[tiles_host_[t].block];
            QZ_HIP(hipMemcpyAsync(lgrp_, lg.data(), nlarge * sizeof(Grp), hipMemcpyHostToDevice, s));
            QZ_HIP(hipMemcpyAsync(ldelta_, delta.data(), nlarge * 4, hipMemcpyHostToDevice, s));
            QZ_HIP(hipMemcpyAsync(tiles_, tiles_host_.data(), lt * sizeof(TileDesc), hipMemcpyHostToDevice, s));
        }
        const TileDesc* ltiles = (const TileDesc*)tiles_;
        // group depths: read from this round's buffer (written when the
        // groups were emitted), new subgroups' depths go to the other one
        const uint32_t* dep_in = rle ? dep_[cur] : nullptr;
        uint32_t* dep_out = rle ? dep_[cur ^ 1] : nullptr;

        // phase A: gather every key of the round before any rank changes
        if (small_waves) {
            hipLaunchKernelGGL(k_gather_small<kGatherTasks>,
                               dim3(qz_div_up(small_waves, (kThreads / QZ_WAVE) * kGatherTasks)), dim3(kThreads),
                               occ_lds(kOccGatherSmall), s, cl, plan, sa_, rank_, bmap_, cb, h, dep_in, bufB_);
            QZ_HIP(hipGetLastError());
        }
        if (nwave) {
            hipLaunchKernelGGL(k_gather_wave, dim3(nwave), dim3(QZ_WAVE), 0, s, cl.list[kClsWave],
                               C.cnt + kClsWave * kCtrStride, sa_, rank_, bmap_, cb, h, dep_in, bufB_);
            QZ_HIP(hipGetLastError());
        }
        if (nmed) {
            hipLaunchKernelGGL(k_gather_med, dim3(nmed), dim3(kThreads), 0, s, cl.list[kClsMed],
                               C.cnt + kClsMed * kCtrStride, sa_, rank_, bmap_, cb, h, dep_in, bufB_);
            QZ_HIP(hipGetLastError());
        }
        if (M) {
            // key range per group, copied back while the small classes sort
            hipLaunchKernelGGL(k_fill_minmax_init, dim3(qz_div_up(nlarge, kThreads)), dim3(kThreads), 0, s, lmm_,
                               nlarge);
            hipLaunchKernelGGL(k_fill_large, dim3(lt), dim3(kThreads), 0, s, ltiles, (const Grp*)lgrp_, sa_, rank_,
                               bmap_, cb, h, dep_in, bufB_, lmm_);
            QZ_HIP(hipGetLastError());
            QZ_HIP(hipMemcpyAsync(lmm_host_, lmm_, (size_t)nlarge * 8, hipMemcpyDeviceToHost, s));
            QZ_HIP(hipEventRecord(lev_, s));
        }
        // phase B: sort groups, update SA / ranks, emit subgroups
        if (small_waves) {
            const uint64_t slots = (uint64_t)small_waves * QZ_WAVE;
            if (ss_lds < 0) {
                ss_lds = tune_pick(kKnobSortSmallLds, slots);
                ss_timed = tune_wants(kKnobSortSmallLds, slots);
            }
            if (ss_timed) {
                QZ_HIP(hipEventRecord(tev_[6], s));
                ss_slots += slots;
            }
            hipLaunchKernelGGL(k_sort_small, dim3(small_sort_blocks), dim3(kThreads), (uint32_t)ss_lds, s, cl, plan,
                               bufB_, sa_, rank_, N, dep_in, dep_out, bmap_, cb);
            QZ_HIP(hipGetLastError());
            if (ss_timed) {
                QZ_HIP(hipEventRecord(tev_[7], s));
                ss_round = true;
            }
        }
        if (nwave) {
            hipLaunchKernelGGL(k_sort_lds<1>, dim3(nwave), dim3(QZ_WAVE), 0, s, cl.list[kClsWave],
                               C.cnt + kClsWave * kCtrStride, bufB_, sa_, rank_, N, dep_in, dep_out, bmap_, cb);
            QZ_HIP(hipGetLastError());
        }
        if (nmed) {
            hipLaunchKernelGGL(k_sort_lds<4>, dim3(nmed), dim3(kThreads), 0, s, cl.list[kClsMed],
                               C.cnt + kClsMed * kCtrStride, bufB_, sa_, rank_, N, dep_in, dep_out, bmap_, cb);
            QZ_HIP(hipGetLastError());
        }
        if (M) {
            // groups whose keys are all equal go on unsplit; the others are
            // sorted at their slots (tiles, offsets and list rebuilt for them)
            QZ_HIP(hipEventSynchronize(lev_));
            lg_split.clear();
            uint64_t kept_elems = 0;
            for (uint32_t i = 0; i < nlarge; ++i) {
                if (lmm_host_[2 * i] == lmm_host_[2 * i + 1])
                    kept_elems += lg[i].size;
                else
                    lg_split.push_back(lg[i]);
            }
            if (trace_) {
                fprintf(stderr, "qzbsc:   large groups unsplit %u of %u (%llu of %u suffixes)\n",
                        nlarge - (uint32_t)lg_split.size(), nlarge, (unsigned long long)kept_elems, M);
                // suffixes of split groups by the bits their key range spans
                uint64_t byb[4] = {0, 0, 0, 0};  // <= 9, <= 18, <= 27, more
                for (uint32_t i = 0; i < nlarge; ++i) {
                    const uint32_t span = lmm_host_[2 * i + 1] - lmm_host_[2 * i];
                    if (!span) continue;
                    const int b = bitlen(span);
                    byb[b <= 9 ? 0 : b <= 18 ? 1 : b <= 27 ? 2 : 3] += lg[i].size;
                }
                fprintf(stderr, "qzbsc:   split by key span <=9 bits %llu <=18 %llu <=27 %llu more %llu\n",
                        (unsigned long long)byb[0], (unsigned long long)byb[1], (unsigned long long)byb[2],
                        (unsigned long long)byb[3]);
            }
            if (lg_split.size() < nlarge) {
                hipLaunchKernelGGL(k_keep_large, dim3(qz_div_up(nlarge, kThreads)), dim3(kThreads), 0, s,
                                   (const Grp*)lgrp_, nlarge, lmm_, N, bmap_, dep_in, dep_out);
                QZ_HIP(hipGetLastError());
                const uint32_t ns = (uint32_t)lg_split.size();
                M = 0;
                lb.resize(ns + 1);
                for (uint32_t i = 0; i < ns; ++i) {
                    lb[i] = M;
                    M += lg_split[i].size;
                }
                lb[ns] = M;
                if (ns) {
                    lhist = build_tiles(lb.data(), ns, tiles_host_.data(), &lt, 1u << lbits);
                    delta.resize(ns);
                    for (uint32_t i = 0; i < ns; ++i) delta[i] = (int32_t)(lg_split[i].start - lb[i]);
                    for (uint32_t t = 0; t < lt; ++t) tiles_host_[t].start += (uint32_t)delta[tiles_host_[t].block];
                    // stream order: k_keep_large has read the full list first
                    QZ_HIP(hipMemcpyAsync(lgrp_, lg_split.data(), ns * sizeof(Grp), hipMemcpyHostToDevice, s));
                    QZ_HIP(hipMemcpyAsync(ldelta_, delta.data(), ns * 4, hipMemcpyHostToDevice, s));
                    QZ_HIP(hipMemcpyAsync(tiles_, tiles_host_.data(), lt * sizeof(TileDesc), hipMemcpyHostToDevice, s));
                }
            }
        }
        if (M) {
            QZ_HIP(hipEventRecord(tev_[2], s));
            bool alt = false;
            // keys at the groups' slots in bufB (beside the other classes'
            // keys), entries in sa; the other side is (keys0, vals0)
            QZ_TRY(seg_sort_bits<uint32_t>(lbits, bufB_, sa_, keys0_, vals0_, ltiles, lt, lhist, lkey_bits, lhist_, tmp_,
                                           s, &alt, VFromArray{sa_}, nullptr, ldelta_));
            QZ_HIP(hipEventRecord(tev_[3], s));
            large_timed = true;
            large_bits = lbits;
            large_n = M;
            const uint32_t* keys = alt ? keys0_ : bufB_;
            const uint32_t* vals = alt ? vals0_ : sa_;
            hipLaunchKernelGGL(k_tile_lasthead, dim3(lt), dim3(kThreads), 0, s, keys, ltiles, bufA_);
            QZ_HIP(hipGetLastError());
            QZ_TRY((device_scan<OpMax, false>(bufA_, bufA_, lt, tmp_, s)));
            hipLaunchKernelGGL(k_final_large, dim3(lt), dim3(kThreads), occ_lds(kOccFinalLarge), s, ltiles,
                               (const Grp*)lgrp_, keys, vals, bufA_, sa_, rank_, N, bmap_, dep_in, dep_out, cb);
            QZ_HIP(hipGetLastError());
        }
        ++rounds_;
        h = (h > 0x40000000u) ? 0x80000000u : h * 2;
        cur ^= 1;
    }
    return 0;
}

int BwtEngine::encode_device(const uint8_t* text, uint8_t* out, uint32_t n, uint32_t nblk, hipStream_t s)
{
    const BlockMap bmap_{off_, btab_, nblk};
    if (offs_host_.size() != (size_t)nblk + 1 || offs_host_.back() != n) return QZ_BAD_PARAMETER;
    if (n > wcap_ || nblk > cap_blk_) return QZ_BAD_PARAMETER;  // the sort workspace holds one chunk
    build_block_tab(offs_host_, n, btab_host_);
    QZ_HIP(hipMemcpyAsync(btab_, btab_host_.data(), btab_host_.size() * 4, hipMemcpyHostToDevice, s));
    QZ_TRY(sort_suffixes(text, n, nblk, s));
    if (cb_) {  // char mode (read back with round 0's counts): a streaming pass
        hipLaunchKernelGGL(k_bwt_out_chars, dim3(qz_div_up(n, kThreads * kOutItems)), dim3(kThreads), 0, s, text,
                           (const PackParams*)ppd_, sa_, rank_, bmap_, n, out);
        QZ_HIP(hipGetLastError());
    } else {
        // the sort is done with its keys: keys0_ (4 bytes per position)
        // holds the 4-bit text when the alphabet allows it
        hipLaunchKernelGGL(k_pack_nib, dim3(qz_div_up(n, kThreads * 8u)), dim3(kThreads), 0, s, text, n,
                           (const PackParams*)ppd_, keys0_, 0u);
        QZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_bwt_out, dim3(qz_div_up(n, kThreads)), dim3(kThreads), occ_lds(kOccBwtOut), s, text,
                           (const uint32_t*)keys0_, (const PackParams*)ppd_, sa_, rank_, bmap_, n, out);
        QZ_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_bwt_index, dim3(nblk), dim3(64), 0, s, rank_, bmap_, idx_);
    QZ_HIP(hipGetLastError());
    return 0;
}

// QZBSC_CLASSES=0 sorts every batch as one, whatever its blocks' alphabets.
static bool classes_enabled()
{
    const char* e = getenv("QZBSC_CLASSES");
    return !(e && e[0] == '0');
}
constexpr uint32_t kSplitMinBytes = 8u << 20;   // smaller batches are never split
constexpr uint64_t kClassMinBytes = 4u << 20;   // smaller classes are sorted together

// Symbols per packed 32-bit key for sigma symbols (as k_make_pack).
static uint32_t packed_k0(uint32_t sigma)
{
    const uint64_t radix = (uint64_t)sigma + 1;
    uint64_t v = 1;
    uint32_t k = 0;
    while (k < 32 && v * radix <= (1ull << 32)) {
        v *= radix;
        ++k;
    }
    return k ? k : 1;
}

// Groups the blocks of the batch at offs by the round-0 packing they would
// get alone: run-length keys (by code width), or packed keys with at most
// 15 symbols (4-bit text, char-mode entries) or more, each by symbols per
// key. Classes under kClassMinBytes, and blocks too small to be sampled,
// are put together. classes is left empty when the batch forms one class.
int BwtEngine::classify(const uint8_t* text, const std::vector<uint32_t>& offs, hipStream_t s,
                        std::vector<std::vector<uint32_t>>& classes)
{
    classes.clear();
    const uint32_t nblk = (uint32_t)offs.size() - 1, n = offs.back();
    build_block_tab(offs, n, btab_host_);
    QZ_HIP(hipMemcpyAsync(btab_, btab_host_.data(), btab_host_.size() * 4, hipMemcpyHostToDevice, s));
    QZ_HIP(hipMemcpyAsync(off_, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, s));
    QZ_HIP(hipMemsetAsync(cls_, 0, (size_t)nblk * kClassWords * 4, s));
    hipLaunchKernelGGL(k_block_class, dim3(qz_div_up(n, (uint64_t)kThreads * kClassSample)), dim3(kThreads), 0, s,
                       text, n, BlockMap{off_, btab_, nblk}, cls_);
    QZ_HIP(hipGetLastError());
    std::vector<uint32_t> h((size_t)nblk * kClassWords);
    QZ_HIP(hipMemcpyAsync(h.data(), cls_, h.size() * 4, hipMemcpyDeviceToHost, s));
    QZ_HIP(hipStreamSynchronize(s));
    const int force = rle_force();
    std::vector<std::pair<uint32_t, uint32_t>> keyed(nblk);  // (class key, block)
    constexpr uint32_t kMisc = 0xffffffffu;
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t* c = &h[(size_t)b * kClassWords];
        uint32_t sigma = 0;
        for (int w = 0; w < 8; ++w) sigma += (uint32_t)__builtin_popcount(c[w]);
        const uint32_t runs = c[8], seen = c[9];
        uint32_t key = kMisc;
        if (sigma) {
            const bool rle = force >= 0 ? force != 0 : (seen >= 64u && (uint64_t)seen >= 3ull * ((uint64_t)runs + 1));
            if (rle)
                key = 0x100u | (sigma <= 31 ? 0u : 1u);  // 3 or 2 run segments per key
            else
                key = (sigma <= 15 ? 0x200u : 0x300u) | packed_k0(sigma);
        }
        keyed[b] = {key, b};
    }
    std::stable_sort(keyed.begin(), keyed.end(),
                     [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
                         return a.first < b.first;
                     });
    std::vector<uint32_t> misc;
    for (size_t i = 0; i < keyed.size();) {
        size_t j = i;
        uint64_t bytes = 0;
        for (; j < keyed.size() && keyed[j].first == keyed[i].first; ++j)
            bytes += offs[keyed[j].second + 1] - offs[keyed[j].second];
        std::vector<uint32_t> blocks;
        for (size_t k = i; k < j; ++k) blocks.push_back(keyed[k].second);
        if (keyed[i].first == kMisc || bytes < kClassMinBytes)
            misc.insert(misc.end(), blocks.begin(), blocks.end());
        else
            classes.push_back(std::move(blocks));
        i = j;
    }
    if (!misc.empty()) {
        std::sort(misc.begin(), misc.end());
        classes.push_back(std::move(misc));
    }
    if (classes.size() <= 1) classes.clear();
    return 0;
}

int BwtEngine::encode_batch(const uint8_t* text, uint8_t* out, const std::vector<uint32_t>& offs, hipStream_t s,
                            std::vector<int32_t>& idx)
{
    const uint32_t nblk = (uint32_t)offs.size() - 1, n = offs.back();
    if (n > cap_ || nblk > cap_blk_) return QZ_BAD_PARAMETER;
    for (uint32_t b = 0; b < nblk; ++b)
        if (offs[b + 1] - offs[b] > wcap_) return QZ_BAD_PARAMETER;  // reserve() sizes the workspace for it
    idx.assign((size_t)nblk * 256, 0);
    const bool trace = getenv("QZBSC_TRACE") != nullptr;
    // One sub-batch: blocks (batch indexes) laid end to end from position
    // base of src, BWTs to the same positions of dst.
    std::vector<int32_t> part;
    auto run_part = [&](const uint8_t* src, uint8_t* dst, const uint32_t* blocks, size_t nb, uint32_t base) -> int {
        std::vector<uint32_t>& lo = offs_host_;
        lo.assign(nb + 1, 0);
        for (size_t i = 0; i < nb; ++i) lo[i + 1] = lo[i] + (offs[blocks[i] + 1] - offs[blocks[i]]);
        QZ_HIP(hipMemcpyAsync(off_, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, s));
        QZ_TRY(encode_device(src + base, dst + base, lo.back(), (uint32_t)nb, s));
        part.resize(nb * 256);
        QZ_HIP(hipMemcpyAsync(part.data(), idx_, part.size() * 4, hipMemcpyDeviceToHost, s));
        QZ_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < nb; ++i)
            std::copy(part.begin() + i * 256, part.begin() + (i + 1) * 256, idx.begin() + (size_t)blocks[i] * 256);
        if (trace)
            fprintf(stderr, "qzbsc: sub-batch of %zu blocks, %u bytes, k0 %u, char mode %u\n", nb, lo.back(), k0_,
                    cb_);
        return 0;
    };
    // Runs of consecutive entries of `blocks` holding at most wcap_ bytes,
    // laid end to end from position base.
    auto run_chunks = [&](const uint8_t* src, uint8_t* dst, const std::vector<uint32_t>& blocks, uint32_t base) -> int {
        for (size_t i = 0; i < blocks.size();) {
            size_t j = i;
            uint64_t bytes = 0;
            while (j < blocks.size() && (j == i || bytes + (offs[blocks[j] + 1] - offs[blocks[j]]) <= wcap_)) {
                bytes += offs[blocks[j] + 1] - offs[blocks[j]];
                ++j;
            }
            QZ_TRY(run_part(src, dst, blocks.data() + i, j - i, base));
            base += (uint32_t)bytes;
            i = j;
        }
        return 0;
    };
    // staging of a class split: T holds the blocks in class order, O their
    // BWTs; T is written while text is read, O is read while out is written
    uint8_t* T = (text == text_) ? out_ : text_;
    uint8_t* O = (T == text_) ? out_ : text_;
    std::vector<std::vector<uint32_t>> classes;
    if (classes_enabled() && nblk >= 2 && n >= kSplitMinBytes && T != text && O != out)
        QZ_TRY(classify(text, offs, s, classes));
    if (classes.empty()) {
        if (n <= wcap_) {  // one sub-batch, in place
            offs_host_ = offs;
            QZ_HIP(hipMemcpyAsync(off_, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, s));
            QZ_TRY(encode_device(text, out, n, nblk, s));
            QZ_HIP(hipMemcpyAsync(idx.data(), idx_, idx.size() * 4, hipMemcpyDeviceToHost, s));
            QZ_HIP(hipStreamSynchronize(s));
            return 0;
        }
        std::vector<uint32_t> all(nblk);
        for (uint32_t b = 0; b < nblk; ++b) all[b] = b;
        return run_chunks(text, out, all, 0);
    }
    // each block's place in T, and the copy runs into T and back out
    std::vector<uint32_t> at(nblk);
    std::vector<CopyRun> runs;
    {
        uint32_t p = 0;
        for (const auto& c : classes)
            for (uint32_t b : c) {
                at[b] = p;
                p += offs[b + 1] - offs[b];
            }
        if (p != n) return QZ_BAD_PARAMETER;  // every block in exactly one class
        for (int dir = 0; dir < 2; ++dir)
            for (uint32_t b = 0; b < nblk; ++b)
                for (uint32_t o = 0, len = offs[b + 1] - offs[b]; o < len; o += kCopyChunk) {
                    const uint32_t l = std::min(kCopyChunk, len - o);
                    runs.push_back(dir == 0 ? CopyRun{at[b] + o, offs[b] + o, l} : CopyRun{offs[b] + o, at[b] + o, l});
                }
    }
    const uint32_t nrun = (uint32_t)runs.size() / 2;
    if (nrun > copy_runs_cap(cap_, cap_blk_)) return QZ_GPU_ERROR;  // cannot happen: blocks within the caps
    QZ_HIP(hipMemcpyAsync(runs_, runs.data(), runs.size() * sizeof(CopyRun), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_copy_runs, dim3(nrun), dim3(kThreads), 0, s, (const CopyRun*)runs_, text, T);
    QZ_HIP(hipGetLastError());
    for (const auto& c : classes) QZ_TRY(run_chunks(T, O, c, at[c[0]]));
    hipLaunchKernelGGL(k_copy_runs, dim3(nrun), dim3(kThreads), 0, s, (const CopyRun*)runs_ + nrun, O, out);
    QZ_HIP(hipGetLastError());
    return 0;
}

}  // namespace qz
