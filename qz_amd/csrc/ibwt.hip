// Inverse BWT on gfx950 by LF-mapping over a batch of blocks.
//
// Replaces libbsc's bsc_bwt_decode (libsais unbwt; third_party, linked via
// crates/qz-lib/build.rs:19-37) on the decode path of
// crates/qz-lib/src/compression/bsc.rs:384-436.
//
// Block b (n bytes, primary p) has n+1 rows; row p holds the sentinel. For a
// real symbol at BWT index u (row = u < p ? u : u+1),
//     LF(row) = 1 + #(symbols < c in the block) + #(c in U[0..u)),
// i.e. the row's position in a stable counting sort of U, shifted by one for
// the sentinel row; LF(p) = 0. Walking LF from row 0 emits T backwards.
//
// Phases:
//   1. block-aligned tiles of 4096 symbols: per-tile symbol histograms,
//      one exclusive scan laid out [block][symbol][tile], then a stable
//      wave-ballot ranking pass that writes one 32-bit entry per row:
//      LF (bits 0..30) | splitter flag (31). The symbol a row emits, L[row],
//      is F[LF(row)], read from a compact per-block table of the F column.
//   2. every S-th row (S = 32, or 64 for batches of 128 MiB and more) is a
//      splitter; each splitter walks LF until the next
//      splitter, recording the segment length and successor.
//   3. list ranking of the splitter lists by pointer jumping gives each
//      segment its distance from the chain start.
//   4. each splitter walks its segment again and writes its symbols.
#include "prims.h"
#include "engine.h"
#include "../../include/qzbsc.h"

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace qz {

constexpr int kIThreads = 256;
constexpr uint32_t kSplitFlag = 1u << 31;  // row entry: LF in bits 0..30, splitter flag in bit 31
constexpr uint32_t kLFMask = kSplitFlag - 1;
constexpr uint32_t kNil = 0xffffffffu;

__global__ __launch_bounds__(kIThreads) void k_ib_hist(const uint8_t* __restrict__ U,
                                                        const TileDesc* __restrict__ tiles,
                                                        uint32_t* __restrict__ hist)
{
    // R copies of each wave's counters, a value's copies in adjacent words,
    // chosen by lane: 16 counts per lane per load from few distinct bytes
    // would otherwise queue on a handful of words
    constexpr uint32_t R = 4;
    __shared__ uint32_t cnt[4][256 * R];
    const TileDesc td = tiles[xcd_item(blockIdx.x, gridDim.x)];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, rep = tid & (R - 1);
    for (uint32_t j = tid; j < 256 * R; j += kIThreads)
        for (int w = 0; w < 4; ++w) cnt[w][j] = 0;
    __syncthreads();
    // tiles start at any byte: the 16-byte aligned middle in 16-byte loads,
    // the head before it and the tail after it byte by byte
    const uint8_t* t = U + td.start;
    const uint32_t head = min((uint32_t)((16u - ((uintptr_t)t & 15u)) & 15u), td.len);
    const uint32_t nv = (td.len - head) / 16u, tail = head + nv * 16u;
    for (uint32_t i = tid; i < head; i += kIThreads) atomicAdd(&cnt[wave][t[i] * R + rep], 1u);
    for (uint32_t i = tail + tid; i < td.len; i += kIThreads) atomicAdd(&cnt[wave][t[i] * R + rep], 1u);
    const uint4* v = reinterpret_cast<const uint4*>(t + head);
    for (uint32_t i = tid; i < nv; i += kIThreads) {
        const uint4 x = v[i];
        const uint32_t w4[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int k = 0; k < 4; ++k)
#pragma unroll
            for (int b = 0; b < 4; ++b) atomicAdd(&cnt[wave][((w4[k] >> (8 * b)) & 255u) * R + rep], 1u);
    }
    __syncthreads();
    uint32_t s = 0;
    for (int w = 0; w < 4; ++w)
#pragma unroll
        for (uint32_t r = 0; r < R; ++r) s += cnt[w][tid * R + r];
    hist[(size_t)td.hbase + (size_t)tid * td.ntiles + td.tlocal] = s;
}

// hist has been exclusive-scanned: entry = off[b] + #(< c in block) + #(c in
// earlier tiles of the block). Each wave ranks its own 1024 consecutive
// symbols (16 steps of 64) against wave-private counters with one ballot per
// symbol bit, so ranking needs no workgroup barrier; one barrier then turns
// the four waves' counts into each wave's base per symbol, and the row
// entries are written. Stable: order within a tile is wave, then step, then
// lane, which is BWT index order.
__global__ __launch_bounds__(kIThreads) void k_ib_rank(const uint8_t* __restrict__ U,
                                                        const TileDesc* __restrict__ tiles,
                                                        const uint32_t* __restrict__ hist,
                                                        const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ prim,
                                                        uint32_t split, uint32_t* __restrict__ E)
{
    constexpr int W = kIThreads / QZ_WAVE;
    constexpr int IT = kSortTile / kIThreads;  // 16 steps of 64 per wave
    __shared__ uint32_t cnt[W][256];
    __shared__ unsigned long long msk[W][256];  // per wave and symbol: the step's lanes holding it
    const TileDesc td = tiles[xcd_item(blockIdx.x, gridDim.x)];
    const uint32_t tid = threadIdx.x, wave = tid / QZ_WAVE, lane = lane_id();
    const uint32_t ob = off[td.block];
    const uint32_t p = prim[td.block];
    const uint32_t rb = ob + td.block;  // row base of this block
    const uint32_t wbase = wave * (IT * QZ_WAVE);
#pragma unroll
    for (int w = 0; w < W; ++w) {
        cnt[w][tid] = 0;
        msk[w][tid] = 0ull;
    }
    // the counters are cleared before any load is issued (a workgroup
    // barrier waits for every load in flight); then all of a lane's symbols
    // go out at once and are ranked as they arrive (lanes past the tile's
    // end read its last symbol and rank nothing)
    __syncthreads();
    const uint32_t last = td.start + td.len - 1u;
    uint8_t cs[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) cs[it] = U[min(td.start + wbase + (uint32_t)it * QZ_WAVE + lane, last)];
    uint32_t rk[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const uint32_t i = wbase + it * QZ_WAVE + lane;
        const bool valid = i < td.len;
        const uint32_t c = cs[it];
        // the lanes of this step holding c: each ORs its bit into the
        // symbol's word and reads the word back; a wave's LDS instructions
        // complete in order, so the read follows every OR of the step, and
        // the symbol's first lane clears the word for the next step
        if (valid) atomicOr(&msk[wave][c], 1ull << lane);
        const uint64_t peer = valid ? (uint64_t)msk[wave][c] : 0ull;
        const uint32_t below = popc_below(peer);
        // a wave's LDS reads and writes complete in order: the next step
        // reads what this one wrote
        const uint32_t before = cnt[wave][c];
        if (valid && below == 0) {
            cnt[wave][c] = before + (uint32_t)__popcll(peer);
            msk[wave][c] = 0ull;
        }
        rk[it] = valid ? before + below : 0xffffffffu;
    }
    __syncthreads();
    {
        // symbol tid: each wave's base = the tile's run + the earlier waves' counts
        uint32_t r = hist[(size_t)td.hbase + (size_t)tid * td.ntiles + td.tlocal] - ob;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            const uint32_t t = cnt[w][tid];
            cnt[w][tid] = r;
            r += t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        if (rk[it] != 0xffffffffu) {
            const uint32_t i = wbase + it * QZ_WAVE + lane;
            const uint32_t u = td.start - ob + i;  // local BWT index
            const uint32_t row = u < p ? u : u + 1u;
            const uint32_t lf = 1u + cnt[wave][cs[it]] + rk[it];  // local row
            E[rb + row] = (row % split == 0) ? (lf | kSplitFlag) : lf;
        }
    }
}

// Sentinel rows: LF(p) = 0.
__global__ void k_ib_sentinel(const uint32_t* __restrict__ off, const uint32_t* __restrict__ prim, uint32_t nblk,
                              uint32_t split, uint32_t* __restrict__ E, uint32_t* __restrict__ ntail)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *ntail = 0u;  // the walks' tail count, before the first walk
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblk) return;
    const uint32_t p = prim[b];
    E[off[b] + b + p] = (p % split == 0) ? kSplitFlag : 0u;
}

// F column of every block, compactly: its present symbols in byte order and
// the local row where each one's run of F starts (1 + #smaller symbols; row 0
// belongs to the sentinel). The symbol emitted at a row is L[row] =
// F[LF(row)], so the walks find it from the LF they already hold. One
// workgroup per block (the block's first tile), one thread per byte value.
__global__ __launch_bounds__(kIThreads) void k_ib_ftable(const TileDesc* __restrict__ tiles,
                                                          const uint32_t* __restrict__ hist,
                                                          const uint32_t* __restrict__ off,
                                                          uint32_t* __restrict__ fst, uint8_t* __restrict__ fsym,
                                                          uint32_t* __restrict__ fsig)
{
    __shared__ uint32_t wsum[kIThreads / QZ_WAVE];
    const TileDesc td = tiles[blockIdx.x];
    if (td.tlocal != 0) return;  // block-uniform
    const uint32_t c = threadIdx.x, lane = lane_id(), w = c / QZ_WAVE;
    const uint32_t b = td.block, ob = off[b], nb = off[b + 1] - ob;
    const uint32_t cum = hist[(size_t)td.hbase + (size_t)c * td.ntiles] - ob;
    const uint32_t nxt = (c + 1 < 256) ? hist[(size_t)td.hbase + (size_t)(c + 1) * td.ntiles] - ob : nb;
    const uint32_t present = nxt > cum ? 1u : 0u;
    uint32_t incl = present;
    for (int d = 1; d < QZ_WAVE; d <<= 1) {
        const uint32_t o = __shfl_up(incl, d, QZ_WAVE);
        if (lane >= (uint32_t)d) incl += o;
    }
    if (lane == QZ_WAVE - 1) wsum[w] = incl;
    __syncthreads();
    for (uint32_t i = 0; i < w; ++i) incl += wsum[i];
    if (present) {
        fst[(size_t)b * 256 + incl - 1] = 1u + cum;
        fsym[(size_t)b * 256 + incl - 1] = (uint8_t)c;
    }
    if (c == kIThreads - 1) fsig[b] = incl;
}

// Symbol whose run of F holds local row x (x >= 1; row 0 gives the first).
// A walking lane holds its block's table in registers when the block has at
// most kRegSyms symbols (DNA, binned qualities): the lookup is then compares
// only. Global loads return in issue order, so a lookup through memory would
// wait behind the walk's own long-latency load; larger alphabets take that
// path (binary search over the table).
constexpr int kRegSyms = 8;
struct FTable {
    uint32_t start[kRegSyms];
    uint32_t syms[kRegSyms / 4];  // four symbols per word, low byte first
    const uint32_t* fst;
    const uint8_t* fsym;
    uint32_t sig;

    __device__ __forceinline__ void load(const uint32_t* __restrict__ fst_all, const uint8_t* __restrict__ fsym_all,
                                         const uint32_t* __restrict__ fsig, uint32_t b)
    {
        fst = fst_all + (size_t)b * 256;
        fsym = fsym_all + (size_t)b * 256;
        sig = fsig[b];
        if (sig <= (uint32_t)kRegSyms) {
#pragma unroll
            for (int k = 0; k < kRegSyms; ++k) start[k] = (uint32_t)k < sig ? fst[k] : 0xffffffffu;
#pragma unroll
            for (int w = 0; w < kRegSyms / 4; ++w) {
                uint32_t v = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if ((uint32_t)(4 * w + q) < sig) v |= (uint32_t)fsym[4 * w + q] << (8 * q);
                syms[w] = v;
            }
        }
    }
    __device__ __forceinline__ uint8_t lookup(uint32_t x) const
    {
        if (sig <= (uint32_t)kRegSyms) {
            uint32_t sym = syms[0] & 0xffu;
#pragma unroll
            for (int k = 1; k < kRegSyms; ++k)
                if (start[k] <= x) sym = (syms[k / 4] >> (8 * (k % 4))) & 0xffu;
            return (uint8_t)sym;
        }
        uint32_t lo = 0, hi = sig;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if (fst[mid] <= x) lo = mid; else hi = mid;
        }
        return fsym[lo];
    }
};

__device__ __forceinline__ uint32_t find_block_u32(const uint32_t* __restrict__ base, uint32_t nblk, uint32_t i)
{
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (base[mid] <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Phase 2: walk from each splitter to the next one. Segment lengths are
// geometric, so lanes pull splitters from a per-workgroup pool in LDS as
// soon as their current segment ends instead of idling until the longest
// lane of the wave is done.
constexpr uint32_t kMinSplit = 32;  // splitter spacing limits (sizes the per-splitter arrays)
constexpr uint32_t kMaxSplit = 256;
constexpr uint64_t kWideSplitBytes = 128ull << 20;  // batches from here on use spacing 64
constexpr uint32_t kPoolMax = 512;  // splitters per walking workgroup (swept: 256..2048)
constexpr uint32_t kPoolLimit = 4096;  // the most a pool may be set to
// A segment keeps its first 2 * split symbols from the first walk, so the
// buffer costs 2 bytes per row whatever the spacing.
__host__ __device__ constexpr uint32_t seg_buf(uint32_t split) { return 2 * split; }
// Buffer bytes for any spacing in [kMinSplit, kMaxSplit]: sum over blocks of
// (rows_b / split + 1) * 2 * split <= 2 * rows + 2 * kMaxSplit * blocks.
static inline uint64_t seg_buf_bytes(uint64_t rows, uint64_t blocks) { return 2 * rows + 2ull * kMaxSplit * blocks + 64; }

// Phase 2 also keeps the first kBuf symbols of every segment (in walk order)
// and, for longer segments, the row reached after kBuf steps.
__global__ __launch_bounds__(kIThreads) void k_ib_walk1(const uint32_t* __restrict__ E,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ sbase, uint32_t nblk,
                                                         const uint32_t* __restrict__ fst,
                                                         const uint8_t* __restrict__ fsym,
                                                         const uint32_t* __restrict__ fsig,
                                                         uint32_t nsplit, uint32_t split, uint32_t pool,
                                                         uint32_t* __restrict__ nxt, uint32_t* __restrict__ len,
                                                         uint8_t* __restrict__ buf, uint32_t* __restrict__ resume,
                                                         uint32_t* __restrict__ tails, uint32_t* __restrict__ ntail)
{
    __shared__ uint32_t next;
    __shared__ uint32_t lt, lbase;
    __shared__ uint32_t ltail[kPoolLimit];  // this pool's splitters with a tail
    const uint32_t kBuf = seg_buf(split);
    const uint32_t pbeg = blockIdx.x * pool;
    const uint32_t pend = min(pbeg + pool, nsplit);
    if (threadIdx.x == 0) {
        next = pbeg;
        lt = 0;
    }
    __syncthreads();
    uint32_t sid = atomicAdd(&next, 1u);
    bool active = sid < pend;
    uint32_t rb = 0, sb = 0, x = 0, l = 0, e = 0;
    FTable ft;
    ft.sig = 0;
    uint64_t word = 0;  // symbols of the current 8-step group, low byte first
    if (active) {
        const uint32_t b = find_block_u32(sbase, nblk, sid);
        rb = off[b] + b;
        sb = sbase[b];
        ft.load(fst, fsym, fsig, b);
        e = E[rb + (sid - sb) * split];
    }
    uint64_t* buf64 = reinterpret_cast<uint64_t*>(buf);
    while (__ballot(active)) {
        if (active) {
            x = e & kLFMask;
            const uint32_t en = E[rb + x];  // issued before the symbol lookup, which it does not need
            if (l < kBuf) {
                word |= (uint64_t)ft.lookup(x) << (8 * (l & 7));
                if ((l & 7) == 7) {
                    buf64[(size_t)sid * (kBuf / 8) + (l >> 3)] = word;
                    word = 0;
                }
            }
            ++l;
            if (l == kBuf) resume[sid] = x;
            if (en & kSplitFlag) {
                if (l < kBuf && (l & 7)) buf64[(size_t)sid * (kBuf / 8) + (l >> 3)] = word;
                word = 0;
                len[sid] = l;
                if (l > kBuf) ltail[atomicAdd(&lt, 1u)] = sid;
                nxt[sid] = (x == 0) ? kNil : sb + x / split;
                sid = atomicAdd(&next, 1u);
                active = sid < pend;
                l = 0;
                if (active) {
                    const uint32_t b = find_block_u32(sbase, nblk, sid);
                    rb = off[b] + b;
                    sb = sbase[b];
                    ft.load(fst, fsym, fsig, b);
                    e = E[rb + (sid - sb) * split];
                }
            } else {
                e = en;
            }
        }
    }
    // the pool's tails go to the batch list, one reservation per workgroup
    __syncthreads();
    if (threadIdx.x == 0) lbase = lt ? atomicAdd(ntail, lt) : 0u;
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < lt; k += kIThreads) tails[lbase + k] = ltail[k];
}

// Phase 3: pointer jumping; dist[s] ends as the row count from s to the end
// of its chain (inclusive).
__global__ __launch_bounds__(kIThreads) void k_ib_jump(const uint32_t* __restrict__ nin,
                                                        const uint32_t* __restrict__ din, uint32_t* __restrict__ nout,
                                                        uint32_t* __restrict__ dout, uint32_t nsplit,
                                                        uint32_t* __restrict__ any)
{
    const uint32_t s = blockIdx.x * kIThreads + threadIdx.x;
    if (s >= nsplit) return;
    const uint32_t n1 = nin[s];
    if (n1 == kNil) {
        nout[s] = kNil;
        dout[s] = din[s];
        return;
    }
    dout[s] = din[s] + din[n1];
    const uint32_t n2 = nin[n1];
    nout[s] = n2;
    if (n2 != kNil) *any = 1u;
}

// Phase 3 by a ruling set: every kRule-th splitter of a block (its first,
// the head of the block's chain, included) is a ruler. Each ruler walks its
// chain to the next ruler, recording for every splitter it passes its ruler
// (owner) and the rows from the ruler to that splitter (pre); rnxt / rlen
// hold the next ruler (kNil at the chain end) and the rows up to it. Pointer
// jumping then ranks the rulers alone, and every splitter's distance follows
// from its ruler's. Rulers are numbered per block: block b's lie at
// rbase[b] .. rbase[b + 1).
constexpr uint32_t kRule = 16;

__global__ __launch_bounds__(kIThreads) void k_ib_rule_walk(const uint32_t* __restrict__ nxt,
                                                             const uint32_t* __restrict__ len,
                                                             const uint32_t* __restrict__ sbase,
                                                             const uint32_t* __restrict__ rbase, uint32_t nblk,
                                                             uint32_t nrule, uint32_t* __restrict__ owner,
                                                             uint32_t* __restrict__ pre, uint32_t* __restrict__ rnxt,
                                                             uint32_t* __restrict__ rlen)
{
    const uint32_t q = blockIdx.x * kIThreads + threadIdx.x;
    if (q >= nrule) return;
    const uint32_t b = find_block_u32(rbase, nblk, q);
    const uint32_t sb = sbase[b], r0 = rbase[b];
    const uint32_t sid = sb + (q - r0) * kRule;
    uint32_t acc = len[sid];
    uint32_t t = nxt[sid];
    // chains stay inside their block, and every chain ends (kNil)
    while (t != kNil && (t - sb) % kRule != 0) {
        owner[t] = q;
        pre[t] = acc;
        acc += len[t];
        t = nxt[t];
    }
    rnxt[q] = (t == kNil) ? kNil : r0 + (t - sb) / kRule;
    rlen[q] = acc;
}

// dist[s]: the rows from splitter s to the end of its chain, from its ruler's.
// A valid block's LF is one cycle through row 0, so every splitter is on its
// block's chain and has a ruler; a BWT whose LF splits into several cycles
// (a corrupt block) leaves some splitters unvisited, and the owner check
// keeps their reads inside the ruler arrays (the writes that follow clamp
// their positions to the block).
__global__ __launch_bounds__(kIThreads) void k_ib_rule_apply(const uint32_t* __restrict__ sbase,
                                                              const uint32_t* __restrict__ rbase, uint32_t nblk,
                                                              uint32_t nsplit, uint32_t nrule,
                                                              const uint32_t* __restrict__ owner,
                                                              const uint32_t* __restrict__ pre,
                                                              const uint32_t* __restrict__ rdist,
                                                              uint32_t* __restrict__ dist)
{
    const uint32_t s = blockIdx.x * kIThreads + threadIdx.x;
    if (s >= nsplit) return;
    const uint32_t b = find_block_u32(sbase, nblk, s);
    const uint32_t local = s - sbase[b];
    if (local % kRule == 0) {
        dist[s] = rdist[rbase[b] + local / kRule];
    } else {
        const uint32_t o = owner[s];
        dist[s] = o < nrule ? rdist[o] - pre[s] : 0u;
    }
}

// Rulers of a batch: block b has its splitters / kRule, rounded up.
static inline uint64_t rule_cap(uint64_t nsplit, uint32_t nblk) { return nsplit / kRule + nblk + 2; }

// Phase 4a: place the buffered symbols, eight per thread: step k of a
// segment starting at chain step t0 writes T[n-1-t0-k]; thread q of a
// splitter takes steps 8q .. 8q+7 from one 8-byte word of the buffer.
__global__ __launch_bounds__(kIThreads) void k_ib_copy(const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ sbase, uint32_t nblk,
                                                        uint32_t nsplit, uint32_t split,
                                                        const uint32_t* __restrict__ len,
                                                        const uint32_t* __restrict__ dist,
                                                        const uint8_t* __restrict__ buf, uint8_t* __restrict__ T)
{
    __shared__ uint32_t b0;
    const uint32_t kChunks = seg_buf(split) / 8;
    const uint64_t gid = (uint64_t)blockIdx.x * kIThreads + threadIdx.x;
    const uint32_t sid = (uint32_t)(gid / kChunks), q = (uint32_t)(gid % kChunks);
    // the block of the workgroup's first splitter, searched once; a thread
    // steps on from it to its own (a workgroup spans few splitters)
    if (threadIdx.x == 0) b0 = find_block_u32(sbase, nblk, (uint32_t)(((uint64_t)blockIdx.x * kIThreads) / kChunks));
    __syncthreads();
    if (sid >= nsplit) return;
    const uint32_t l = len[sid];
    const uint32_t ds = dist[sid];
    const uint64_t w = reinterpret_cast<const uint64_t*>(buf)[(size_t)sid * kChunks + q];
    const uint32_t k0 = q * 8;
    if (k0 >= l) return;
    uint32_t b = b0;
    while (b + 1 < nblk && sbase[b + 1] <= sid) ++b;
    const uint32_t ob = off[b], n = off[b + 1] - ob;
    const uint32_t t0 = n + 1u - ds + k0;
    const uint32_t m = l - k0 < 8u ? l - k0 : 8u;
    for (uint32_t j = 0; j < m; ++j) {
        const uint32_t t = t0 + j;
        if (t < n) T[ob + n - 1u - t] = (uint8_t)(w >> (8 * j));
    }
}

// Phase 4b: segments longer than kBuf walk on from their resume row.
__global__ __launch_bounds__(kIThreads) void k_ib_walk2(const uint32_t* __restrict__ E,
                                                         const uint32_t* __restrict__ off,
                                                         const uint32_t* __restrict__ sbase, uint32_t nblk,
                                                         const uint32_t* __restrict__ fst,
                                                         const uint8_t* __restrict__ fsym,
                                                         const uint32_t* __restrict__ fsig,
                                                         uint32_t split, uint32_t pool,
                                                         const uint32_t* __restrict__ tails,
                                                         const uint32_t* __restrict__ ntail,
                                                         const uint32_t* __restrict__ len,
                                                         const uint32_t* __restrict__ dist,
                                                         const uint32_t* __restrict__ resume, uint8_t* __restrict__ T)
{
    // the list of tails (walk1 appends every splitter whose segment is
    // longer than its buffer) in pools of `pool` per workgroup; the grid
    // covers the longest possible list, and workgroups past its end return
    __shared__ uint32_t next;
    const uint32_t kBuf = seg_buf(split);
    const uint32_t pbeg = blockIdx.x * pool;
    const uint32_t pend = min(pbeg + pool, *ntail);
    if (pbeg >= pend) return;  // uniform
    if (threadIdx.x == 0) next = pbeg;
    __syncthreads();
    uint32_t rb = 0, l = 0, t = 0, last = 0, e = 0;
    FTable ft;
    ft.sig = 0;
    uint8_t* dst = nullptr;
    auto pull = [&]() -> bool {
        const uint32_t k = atomicAdd(&next, 1u);
        if (k >= pend) return false;
        const uint32_t sid = tails[k];
        const uint32_t b = find_block_u32(sbase, nblk, sid);
        const uint32_t ob = off[b], n = off[b + 1] - ob;
        rb = ob + b;
        l = len[sid] - kBuf;
        t = n + 1u - dist[sid] + kBuf;
        last = n;
        dst = T + ob + n - 1u;  // dst - t addresses T[n-1-t]
        ft.load(fst, fsym, fsig, b);
        e = E[rb + resume[sid]];
        return true;
    };
    // A lane writes its tail downwards, one byte per step. It gathers them
    // per aligned 8-byte window of T and stores a window it fills whole with
    // one 8-byte store (a wave's byte stores would each be a request of
    // their own); the partial windows at a tail's two ends, whose other
    // bytes belong to other segments, go out byte by byte.
    uint64_t acc = 0;
    uint32_t have = 0;      // bytes of the window gathered so far (bit k: offset k)
    uint8_t* win = nullptr;  // the window's first byte
    auto flush = [&]() {
        if (have == 0xffu) {
            *reinterpret_cast<uint64_t*>(win) = acc;
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k)
                if ((have >> k) & 1u) win[k] = (uint8_t)(acc >> (8 * k));
        }
        acc = 0;
        have = 0;
    };
    bool active = pull();
    while (__ballot(active)) {
        if (active) {
            const uint32_t x = e & kLFMask;
            if (--l != 0) e = E[rb + x];  // issued before the symbol lookup
            if (t < last) {
                uint8_t* const q = dst - (int64_t)t;
                const uint32_t o = (uint32_t)((uintptr_t)q & 7u);
                win = q - o;
                acc |= (uint64_t)ft.lookup(x) << (8 * o);
                have |= 1u << o;
                if (o == 0) flush();  // the window's lowest byte is its last step
            }
            ++t;
            if (l == 0) {
                if (have) flush();
                active = pull();
            }
        }
    }
}

uint64_t IbwtEngine::footprint(uint64_t n, uint32_t nblk)
{
    const uint64_t c = std::max<uint64_t>(n, 1024);
    const uint32_t cb = std::max<uint32_t>(nblk, 1);
    const uint64_t rows = c + cb, tiles = c / kSortTile + cb + 1, nsplit = rows / kMinSplit + cb + 1;
    return 2 * (c + 64) + rows * 4 + (uint64_t)cb * (256 * 5 + 4) + tiles * 256 * 4 + (scan_tmp_words(tiles * 256) + 64) * 4 +
           tiles * sizeof(TileDesc) + 4 * (cb + 1) * 4 + 7 * nsplit * 4 + 5 * rule_cap(nsplit, cb) * 4 +
           seg_buf_bytes(rows, cb) + 64;
}

int IbwtEngine::reserve(uint64_t n, uint32_t nblk)
{
    if (n <= cap_ && nblk <= cap_blk_) return 0;
    const MemTrace mt("inverse", n, nblk);
    // grow geometrically so a stream of slowly growing batches reallocates
    // only a few times
    const uint64_t c = std::max<uint64_t>(std::max<uint64_t>(n, 1024), cap_ ? cap_ + cap_ / 2 : 0);
    const uint32_t cb = std::max<uint32_t>(std::max<uint32_t>(nblk, 1), cap_blk_ ? cap_blk_ + cap_blk_ / 2 : 0);
    release();
    const uint64_t rows = c + cb;
    const uint64_t tiles = c / kSortTile + cb + 1;
    const uint64_t nsplit = rows / kMinSplit + cb + 1;
    QZ_HIP(hipMalloc(&U_, c + 64));
    QZ_HIP(hipMalloc(&T_, c + 64));
    QZ_HIP(hipMalloc(&E_, rows * 4));
    QZ_HIP(hipMalloc(&fst_, (size_t)cb * 256 * 4));
    QZ_HIP(hipMalloc(&fsym_, (size_t)cb * 256));
    QZ_HIP(hipMalloc(&fsig_, (size_t)cb * 4));
    QZ_HIP(hipMalloc(&hist_, tiles * 256 * 4));
    QZ_HIP(hipMalloc(&tmp_, (scan_tmp_words(tiles * 256) + 64) * 4));
    QZ_HIP(hipMalloc(&tiles_, tiles * sizeof(TileDesc)));
    QZ_HIP(hipMalloc(&off_, (cb + 1) * 4));
    QZ_HIP(hipMalloc(&sbase_, (cb + 1) * 4));
    QZ_HIP(hipMalloc(&rbase_, (cb + 1) * 4));
    QZ_HIP(hipMalloc(&rul_, 5 * rule_cap(nsplit, cb) * 4));
    QZ_HIP(hipMalloc(&prim_, cb * 4));
    QZ_HIP(hipMalloc(&nxt0_, nsplit * 4));
    QZ_HIP(hipMalloc(&nxt1_, nsplit * 4));
    QZ_HIP(hipMalloc(&len_, nsplit * 4));
    QZ_HIP(hipMalloc(&d0_, nsplit * 4));
    QZ_HIP(hipMalloc(&d1_, nsplit * 4));
    QZ_HIP(hipMalloc(&flag_, 64));
    QZ_HIP(hipMalloc(&buf_, seg_buf_bytes(rows, cb)));
    QZ_HIP(hipMalloc(&resume_, nsplit * 4));
    QZ_HIP(hipMalloc(&tails_, nsplit * 4));
    QZ_HIP(hipHostMalloc(&pinned_, 64));
    cap_ = c;
    cap_blk_ = cb;
    cap_tiles_ = tiles;
    cap_split_ = nsplit;
    return 0;
}

// The workspace of bind(): the same arrays reserve() allocates, but U_/T_,
// carved from one region at 256-byte boundaries. With base == nullptr only
// sizes the region (returns its bytes).
static size_t carve(uint8_t* base, uint64_t n, uint32_t nblk, uint32_t** E, uint32_t** fst, uint8_t** fsym,
                    uint32_t** fsig, uint32_t** hist, uint32_t** tmp, void** tiles, uint32_t** off, uint32_t** sbase,
                    uint32_t** prim, uint32_t** nxt0, uint32_t** nxt1, uint32_t** len, uint32_t** d0, uint32_t** d1,
                    uint32_t** flag, uint8_t** buf, uint32_t** resume, uint32_t** rbase, uint32_t** rul,
                    uint32_t** tails)
{
    const uint64_t c = std::max<uint64_t>(n, 1024);
    const uint32_t cb = std::max<uint32_t>(nblk, 1);
    const uint64_t rows = c + cb, tiles_n = c / kSortTile + cb + 1, nsplit = rows / kMinSplit + cb + 1;
    size_t at = 0;
    auto take = [&](auto** p, uint64_t bytes) {
        if (base) *p = reinterpret_cast<std::remove_pointer_t<decltype(p)>>(base + at);
        at += (bytes + 255) & ~(uint64_t)255;
    };
    take(E, rows * 4);
    take(fst, (uint64_t)cb * 256 * 4);
    take(fsym, (uint64_t)cb * 256);
    take(fsig, (uint64_t)cb * 4);
    take(hist, tiles_n * 256 * 4);
    take(tmp, (scan_tmp_words(tiles_n * 256) + 64) * 4);
    take(tiles, tiles_n * sizeof(TileDesc));
    take(off, (uint64_t)(cb + 1) * 4);
    take(sbase, (uint64_t)(cb + 1) * 4);
    take(prim, (uint64_t)cb * 4);
    take(nxt0, nsplit * 4);
    take(nxt1, nsplit * 4);
    take(len, nsplit * 4);
    take(d0, nsplit * 4);
    take(d1, nsplit * 4);
    take(flag, 64);
    take(buf, seg_buf_bytes(rows, cb));
    take(resume, nsplit * 4);
    take(rbase, (uint64_t)(cb + 1) * 4);
    take(rul, 5 * rule_cap(nsplit, cb) * 4);
    take(tails, nsplit * 4);
    return at;
}

size_t IbwtEngine::bound_bytes(uint64_t n, uint32_t nblk)
{
    uint32_t *a, *b, *c, *d, *e, *f, *g, *h, *i, *j, *k, *l, *m, *o, *q, *r, *u, *v;
    uint8_t *y, *z;
    void* t;
    return carve(nullptr, n, nblk, &a, &b, &y, &c, &d, &e, &t, &f, &g, &h, &i, &j, &k, &l, &m, &o, &z, &q, &r, &u, &v);
}

int IbwtEngine::bind(uint8_t* base, size_t bytes, uint64_t n, uint32_t nblk)
{
    if (!base || bound_bytes(n, nblk) > bytes) return QZ_NOT_ENOUGH_MEMORY;
    release();
    carve(base, n, nblk, &E_, &fst_, &fsym_, &fsig_, &hist_, &tmp_, &tiles_, &off_, &sbase_, &prim_, &nxt0_, &nxt1_,
          &len_, &d0_, &d1_, &flag_, &buf_, &resume_, &rbase_, &rul_, &tails_);
    const uint64_t c = std::max<uint64_t>(n, 1024);
    const uint32_t cb = std::max<uint32_t>(nblk, 1);
    cap_ = c;
    cap_blk_ = cb;
    cap_tiles_ = c / kSortTile + cb + 1;
    cap_split_ = (c + cb) / kMinSplit + cb + 1;
    borrowed_ = true;
    // host staging at its largest now, so later calls do not grow it
    td_host_.reserve(cap_tiles_);
    sb_host_.reserve((size_t)cb + 1);
    return 0;
}

void IbwtEngine::release()
{
    void* ps[] = {U_,   T_,   E_,    hist_, tmp_, tiles_,  off_, sbase_, prim_, nxt0_, nxt1_,
                  len_, d0_, d1_, flag_, buf_, resume_, fst_, fsym_, fsig_, rbase_, rul_, tails_};
    if (!borrowed_)
        for (void* p : ps)
            if (p) (void)hipFree(p);
    borrowed_ = false;
    if (pinned_) (void)hipHostFree(pinned_);
    U_ = T_ = nullptr;
    E_ = nullptr;
    fst_ = fsig_ = nullptr;
    fsym_ = nullptr;
    hist_ = tmp_ = off_ = sbase_ = prim_ = nxt0_ = nxt1_ = len_ = d0_ = d1_ = flag_ = resume_ = nullptr;
    rbase_ = rul_ = tails_ = nullptr;
    buf_ = nullptr;
    tiles_ = nullptr;
    pinned_ = nullptr;
    cap_ = 0;
    cap_blk_ = 0;
    graph_.reset();
}

// off_ and prim_ are staged; reads U, writes T (device).
int IbwtEngine::decode_device(const uint8_t* U, uint8_t* T, uint32_t n, uint32_t nblk, const std::vector<uint32_t>& offs,
                              hipStream_t s)
{
    // Splitter spacing: longer segments mean fewer splitters to rank, but a
    // batch needs enough segments in flight to hide the walks' latency.
    uint32_t split = split_ ? split_ : (n >= kWideSplitBytes ? 64u : 32u);
    if (const char* e = getenv("QZBSC_SPLIT")) {
        const long v = strtol(e, nullptr, 10);
        if (v >= kMinSplit && v <= kMaxSplit) split = (uint32_t)v;
    }
    uint32_t pool = kPoolMax;  // splitters per walking workgroup
    if (const char* e = getenv("QZBSC_POOL")) {
        const long v = strtol(e, nullptr, 10);
        if (v >= kIThreads && v <= (long)kPoolLimit) pool = (uint32_t)v;
    }
    // block-aligned tiles and per-block splitter bases (sizes only; the
    // descriptors are built when the graph is not replayed)
    uint32_t ntiles = 0;
    const uint32_t hbase = count_tiles(offs.data(), nblk, &ntiles, 256);
    uint32_t nsplit = 0, nrule = 0, chain = 1;  // chain: the most rulers of one block
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t m = (offs[b + 1] - offs[b]) / split + 1;
        const uint32_t r = (m + kRule - 1) / kRule;
        nsplit += m;
        nrule += r;
        chain = r > chain ? r : chain;
    }
    if (ntiles > cap_tiles_ || nsplit > cap_split_) return QZ_BAD_PARAMETER;
    auto prepare = [&](hipStream_t st) -> int {
        // host staging lives in the engine, so the copies may complete late
        td_host_.resize(ntiles);
        uint32_t nt = 0;
        build_tiles(offs.data(), nblk, td_host_.data(), &nt, 256);
        sb_host_.resize(nblk + 1);
        uint32_t acc = 0;
        for (uint32_t b = 0; b < nblk; ++b) {
            sb_host_[b] = acc;
            acc += (offs[b + 1] - offs[b]) / split + 1;
        }
        sb_host_[nblk] = acc;
        rb_host_.resize(nblk + 1);
        acc = 0;
        for (uint32_t b = 0; b < nblk; ++b) {
            rb_host_[b] = acc;
            acc += ((offs[b + 1] - offs[b]) / split + 1 + kRule - 1) / kRule;
        }
        rb_host_[nblk] = acc;
        QZ_HIP(hipMemcpyAsync(tiles_, td_host_.data(), (size_t)nt * sizeof(TileDesc), hipMemcpyHostToDevice, st));
        QZ_HIP(hipMemcpyAsync(sbase_, sb_host_.data(), sb_host_.size() * 4, hipMemcpyHostToDevice, st));
        QZ_HIP(hipMemcpyAsync(rbase_, rb_host_.data(), rb_host_.size() * 4, hipMemcpyHostToDevice, st));
        return 0;
    };
    auto launch = [&](hipStream_t st) -> int {
        hipLaunchKernelGGL(k_ib_hist, dim3(ntiles), dim3(kIThreads), 0, st, U, (const TileDesc*)tiles_, hist_);
        QZ_HIP(hipGetLastError());
        QZ_TRY((device_scan<OpAdd, false>(hist_, hist_, hbase, tmp_, st)));
        hipLaunchKernelGGL(k_ib_rank, dim3(ntiles), dim3(kIThreads), 0, st, U, (const TileDesc*)tiles_, hist_, off_,
                           prim_, split, E_);
        QZ_HIP(hipGetLastError());
        uint32_t* const ntail = flag_ + 1;  // flag_[0]: the jumps' flag
        hipLaunchKernelGGL(k_ib_sentinel, dim3(qz_div_up(nblk, 64)), dim3(64), 0, st, off_, prim_, nblk, split, E_,
                           ntail);
        QZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ib_ftable, dim3(ntiles), dim3(kIThreads), 0, st, (const TileDesc*)tiles_, hist_, off_,
                           fst_, fsym_, fsig_);
        QZ_HIP(hipGetLastError());
        const uint32_t gs = qz_div_up(nsplit, kIThreads);
        const uint32_t gp = qz_div_up(nsplit, pool);
        hipLaunchKernelGGL(k_ib_walk1, dim3(gp), dim3(kIThreads), 0, st, E_, off_, sbase_, nblk, fst_, fsym_, fsig_,
                           nsplit, split, pool, nxt0_, len_, buf_, resume_, tails_, ntail);
        QZ_HIP(hipGetLastError());
        // list ranking by a ruling set: rulers walk to the next ruler
        // (owner / pre in nxt1_ / d1_), pointer jumping ranks the rulers,
        // then every splitter's distance goes to d0_
        const uint64_t rc = rule_cap(cap_split_, cap_blk_);
        uint32_t *rA = rul_, *rB = rul_ + rc, *rL = rul_ + 2 * rc, *rD1 = rul_ + 3 * rc, *rD2 = rul_ + 4 * rc;
        const uint32_t gr = qz_div_up(nrule, kIThreads);
        hipLaunchKernelGGL(k_ib_rule_walk, dim3(gr), dim3(kIThreads), 0, st, nxt0_, len_, sbase_, rbase_, nblk, nrule,
                           nxt1_, d1_, rA, rL);
        QZ_HIP(hipGetLastError());
        // a block's rulers form one chain (its LF cycle), so
        // ceil(log2(longest chain)) + 1 rounds always suffice
        const uint32_t* din = rL;
        uint32_t *na = rA, *nb_ = rB, *dout = rD1;
        int rounds = 1;
        while ((1u << rounds) < chain && rounds < 31) ++rounds;
        for (int r = 0; r <= rounds; ++r) {
            hipLaunchKernelGGL(k_ib_jump, dim3(gr), dim3(kIThreads), 0, st, na, din, nb_, dout, nrule, flag_);
            QZ_HIP(hipGetLastError());
            std::swap(na, nb_);
            din = dout;
            dout = (dout == rD1) ? rD2 : rD1;
        }
        hipLaunchKernelGGL(k_ib_rule_apply, dim3(gs), dim3(kIThreads), 0, st, sbase_, rbase_, nblk, nsplit, nrule,
                           nxt1_, d1_, din, d0_);
        QZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ib_copy, dim3(qz_div_up((uint64_t)nsplit * (seg_buf(split) / 8), kIThreads)),
                           dim3(kIThreads), 0, st, off_, sbase_, nblk, nsplit, split, len_, d0_, buf_, T);
        QZ_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_ib_walk2, dim3(gp), dim3(kIThreads), 0, st, E_, off_, sbase_, nblk, fst_, fsym_, fsig_,
                           split, pool, (const uint32_t*)tails_, (const uint32_t*)ntail, len_, d0_, resume_, T);
        QZ_HIP(hipGetLastError());
        return 0;
    };
    if (borrowed_) {  // bound workspace: eager launches only (capturing a graph could allocate)
        QZ_TRY(prepare(s));
        return launch(s);
    }
    std::vector<uint64_t> key = {n, nblk, split, pool, (uint64_t)(uintptr_t)U, (uint64_t)(uintptr_t)T, cap_, cap_blk_};
    key.insert(key.end(), offs.begin(), offs.end());
    return graph_.run(key, s, prepare, launch);
}

// ---- the inverse BWT without allocation ---------------------------------
//
// For callers that cannot allocate device memory (code under graph capture,
// tight serving loops): the workspace of batches up to kNoAllocBytes input
// bytes in up to kNoAllocBlocks blocks lives in the file-scope array below,
// which every device the library runs on holds until the process ends. Its
// device address is taken once per device with hipGetSymbolAddress, and only
// that address is handed to the engine and its kernels. Every word the
// kernels read within a call is written earlier in the same call, so nothing
// has to be cleared between calls. One device's array is shared by all calls
// on that device: calls are serialised by a per-device lock, and calls that
// overlap from another process or library instance on the same device would
// race.
constexpr uint64_t kNoAllocBytes = 8ull << 20;
constexpr uint32_t kNoAllocBlocks = 64;
constexpr size_t kScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    std::mutex mu;
    bool ready = false;
    uint8_t* ptr = nullptr;  // this device's address of `scratch`
    IbwtEngine eng;
    std::vector<uint32_t> offs;
};
NoAllocDevice g_noalloc[kMaxDevices];
}  // namespace

static int noalloc_setup(NoAllocDevice& D)
{
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        fprintf(stderr, "qzbsc: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
        return QZ_GPU_ERROR;
    }
    D.ptr = (uint8_t*)p;
    QZ_TRY(D.eng.bind(D.ptr, kScratchBytes, kNoAllocBytes, kNoAllocBlocks));
    D.offs.reserve(kNoAllocBlocks + 1);
    D.ready = true;
    return 0;
}

}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API int qzbsc_dev_bwt_decode_batch_noalloc(const void* d_in, void* d_out, const int* sizes, int nblocks,
                                                 const int* primary, void* stream)
{
    static_assert(kNoAllocBlocks <= 256, "block count fits the bound");
    if (!d_in || !d_out || !sizes || !primary || nblocks < 1) return QZ_BAD_PARAMETER;
    if ((uint32_t)nblocks > kNoAllocBlocks) return QZ_NOT_ENOUGH_MEMORY;
    uint64_t tot = 0;
    for (int b = 0; b < nblocks; ++b) {
        if (sizes[b] < 1 || primary[b] < 1 || primary[b] > sizes[b]) return QZ_BAD_PARAMETER;
        tot += (uint64_t)sizes[b];
    }
    if (tot > kNoAllocBytes) return QZ_NOT_ENOUGH_MEMORY;
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return QZ_GPU_ERROR;
    if (dev >= kMaxDevices) return QZ_NOT_SUPPORTED;
    NoAllocDevice& D = g_noalloc[dev];
    std::lock_guard<std::mutex> lk(D.mu);
    if (!D.ready) QZ_TRY(noalloc_setup(D));
    D.offs.resize((size_t)nblocks + 1);
    D.offs[0] = 0;
    for (int b = 0; b < nblocks; ++b) D.offs[b + 1] = D.offs[b] + (uint32_t)sizes[b];
    const hipStream_t s = (hipStream_t)stream;
    QZ_HIP(hipMemcpyAsync(D.eng.off_, D.offs.data(), D.offs.size() * 4, hipMemcpyHostToDevice, s));
    QZ_HIP(hipMemcpyAsync(D.eng.prim_, primary, (size_t)nblocks * 4, hipMemcpyHostToDevice, s));
    QZ_TRY(D.eng.decode_device((const uint8_t*)d_in, (uint8_t*)d_out, (uint32_t)tot, (uint32_t)nblocks, D.offs, s));
    QZ_HIP(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"
