// QLFC front end on gfx950: coder-segment split and the run/rank transform
// of every segment of a batch of BWT blocks.
//
// Replaces the front half of libbsc's qlfc.cpp / coder.cpp (absent from the
// reference tree; built by crates/qz-lib/build.rs:19-37) that bsc_compress
// runs after the BWT (reached from crates/qz-lib/src/compression/bsc.rs:
// 172-183).
//
// Segment split (libbsc coder): with nb segments and `changes` the number
// of i = 1, 33, 65, ... where U[i] != U[i-1], a boundary falls at the
// sample of the (m * (changes / nb))-th change, m = 1 .. nb-1; blocks with
// changes <= nb split evenly.
//
// Transform: maximal runs; symbols kept in a list ordered by next
// occurrence; scanning runs from last to first, a run's rank is its
// symbol's list position (list length if absent), then the symbol moves to
// the front. Parallel form:
//   1. run heads (symbol change or segment start) -> run starts, symbols,
//      lengths by one scan;
//   2. runs are cut into chunks of kChunk runs inside each segment; each
//      chunk's signature = its symbols in first-occurrence order;
//   3. per segment, right to left, the list a chunk starts from is
//      sig(k+1) followed by the rest of chunk k+1's starting list (a
//      two-level scan: stretches of chunks per lane, then across lanes);
//   4. each chunk replays its runs backwards from that list (in LDS).
#include "prims.h"
#include "qlfc_gpu.h"
#include "qlfc_coder.h"

#include <algorithm>
#include <vector>

namespace qz {

constexpr int kFThreads = 256;
constexpr uint32_t kChunk = 512;    // runs per chunk
constexpr int kRankThreads = 64;    // one chunk per thread, 256-byte list each in LDS

static int num_segments_host(int n) { return qlfc::num_segments(n); }

// ---- segment split ---------------------------------------------------------
struct SplitBlock {
    uint32_t off;     // block start
    uint32_t n;       // block size
    uint32_t sbase;   // first sample index (global)
    uint32_t ns;      // samples: i = 1 + 32k < n
    uint32_t nb;      // segments
};

__global__ __launch_bounds__(kFThreads) void k_split_flags(const uint8_t* __restrict__ U,
                                                          const SplitBlock* __restrict__ blks, uint32_t nblk,
                                                          uint32_t nsamples, uint32_t* __restrict__ flags)
{
    const uint32_t k = blockIdx.x * kFThreads + threadIdx.x;
    if (k >= nsamples) return;
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (blks[mid].sbase <= k) lo = mid; else hi = mid;
    }
    const SplitBlock sb = blks[lo];
    const uint32_t i = sb.off + 1 + 32 * (k - sb.sbase);
    flags[k] = U[i] != U[i - 1] ? 1u : 0u;
}

// cum = inclusive scan of the flags. Writes seg starts (block-local) for
// boundaries m = 1 .. nb-1 of blocks with changes > nb.
__global__ __launch_bounds__(kFThreads) void k_split_find(const SplitBlock* __restrict__ blks, uint32_t nblk,
                                                         uint32_t nsamples, const uint32_t* __restrict__ flags,
                                                         const uint32_t* __restrict__ cum,
                                                         uint32_t* __restrict__ seg_start /* nblk x 8 */)
{
    const uint32_t k = blockIdx.x * kFThreads + threadIdx.x;
    if (k >= nsamples || !flags[k]) return;
    uint32_t lo = 0, hi = nblk;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (blks[mid].sbase <= k) lo = mid; else hi = mid;
    }
    const SplitBlock sb = blks[lo];
    const uint32_t before = sb.sbase ? cum[sb.sbase - 1] : 0u;
    const uint32_t total = cum[sb.sbase + sb.ns - 1] - before;
    if (sb.nb < 2 || total <= sb.nb) return;
    const uint32_t per = total / sb.nb;
    const uint32_t c = cum[k] - before;
    if (c % per == 0) {
        const uint32_t m = c / per;
        if (m >= 1 && m < sb.nb) seg_start[lo * 8 + m] = 1 + 32 * (k - sb.sbase);
    }
}

// ---- run detection ---------------------------------------------------------
// head[i] for i in [0, N): symbol change or segment start (segflag set).
__global__ __launch_bounds__(kFThreads) void k_run_heads(const uint8_t* __restrict__ U, uint32_t N,
                                                        const uint8_t* __restrict__ segflag, uint32_t* __restrict__ head)
{
    const uint32_t i = blockIdx.x * kFThreads + threadIdx.x;
    if (i >= N) return;
    head[i] = (i == 0 || segflag[i] || U[i] != U[i - 1]) ? 1u : 0u;
}

__global__ __launch_bounds__(kFThreads) void k_run_emit(const uint8_t* __restrict__ U, uint32_t N,
                                                       const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ ridx,
                                                       uint32_t* __restrict__ rstart, uint8_t* __restrict__ rsym)
{
    const uint32_t i = blockIdx.x * kFThreads + threadIdx.x;
    if (i >= N || !head[i]) return;
    const uint32_t r = ridx[i] - 1;
    rstart[r] = i;
    rsym[r] = U[i];
}

__global__ __launch_bounds__(kFThreads) void k_run_len(const uint32_t* __restrict__ rstart, uint32_t R, uint32_t N,
                                                      uint32_t* __restrict__ rlen)
{
    const uint32_t r = blockIdx.x * kFThreads + threadIdx.x;
    if (r >= R) return;
    rlen[r] = (r + 1 < R ? rstart[r + 1] : N) - rstart[r];
}

// Marks segment starts; gathers scan values at given positions.
__global__ void k_mark(const uint32_t* __restrict__ pos, uint32_t cnt, uint8_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) flag[pos[i]] = 1;
}

__global__ void k_gather(const uint32_t* __restrict__ src, const uint32_t* __restrict__ pos, uint32_t cnt,
                         uint32_t* __restrict__ dst)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cnt) dst[i] = src[pos[i]];
}

// ---- chunk signatures --------------------------------------------------------
struct Chunk {
    uint32_t r0, r1;  // runs [r0, r1)
    uint32_t seg;     // segment index (global)
    uint32_t last;    // 1 if the last chunk of its segment
};

__global__ __launch_bounds__(kFThreads) void k_chunk_sig(const uint8_t* __restrict__ rsym,
                                                        const Chunk* __restrict__ chunks, uint32_t nchunks,
                                                        uint8_t* __restrict__ sig, uint16_t* __restrict__ siglen)
{
    const uint32_t c = blockIdx.x * kFThreads + threadIdx.x;
    if (c >= nchunks) return;
    const Chunk ch = chunks[c];
    uint32_t seen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t L = 0;
    uint8_t* out = sig + (size_t)c * 256;
    for (uint32_t r = ch.r0; r < ch.r1 && L < 256; ++r) {
        const uint32_t s = rsym[r];
        const uint32_t w = s >> 5, bit = 1u << (s & 31);
        uint32_t sw = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q)
            if ((uint32_t)q == w) sw = seen[q];
        if (!(sw & bit)) {
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if ((uint32_t)q == w) seen[q] |= bit;
            out[L++] = (uint8_t)s;
        }
    }
    siglen[c] = (uint16_t)L;
}

// out = sg ++ (L \ sg): the list after a chunk whose signature is sg (its
// symbols in first-occurrence order) is replayed backwards from list L.
// Returns the length of out (out must not alias either input).
__device__ uint32_t prepend_sig(const uint8_t* sg, uint32_t ls, const uint8_t* L, uint32_t ll, uint8_t* out)
{
    uint32_t in[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t q = 0; q < ls; ++q) {
        const uint32_t x = sg[q];
        out[q] = (uint8_t)x;
#pragma unroll
        for (int w = 0; w < 8; ++w)
            if ((uint32_t)w == (x >> 5)) in[w] |= 1u << (x & 31);
    }
    uint32_t lo = ls;
    for (uint32_t q = 0; q < ll; ++q) {
        const uint32_t x = L[q];
        uint32_t iw = 0;
#pragma unroll
        for (int w = 0; w < 8; ++w)
            if ((uint32_t)w == (x >> 5)) iw = in[w];
        if (!(iw & (1u << (x & 31)))) out[lo++] = (uint8_t)x;
    }
    return lo;
}

// One 64-lane workgroup per segment: the list each chunk starts its backward
// replay from (state), and the segment's symbol table. Chunk c maps a list L
// to sig(c) ++ (L \ sig(c)); a run of such maps is again of that form, its
// list being the maps applied to the empty list. So each lane composes one
// stretch of consecutive chunks (A), lane 0 chains the stretches right to
// left into the list entering each stretch (B), and each lane then walks its
// stretch again, writing every chunk's starting list (C).
__global__ __launch_bounds__(64) void k_chunk_state(const uint32_t* __restrict__ seg_chunk0,
                                                    const uint8_t* __restrict__ sig,
                                                    const uint16_t* __restrict__ siglen, uint8_t* __restrict__ state,
                                                    uint16_t* __restrict__ statelen, uint8_t* __restrict__ table,
                                                    uint16_t* __restrict__ nsym)
{
    __shared__ uint8_t comp[64][256];  // (A) stretch composite; (C) work list
    __shared__ uint8_t work[64][256];  // (A), (C) work list
    __shared__ uint8_t inl[64][256];   // (B) list entering each stretch from the right
    __shared__ uint16_t clen[64], ilen[64];
    const uint32_t sgi = blockIdx.x;
    const uint32_t l = threadIdx.x;
    const uint32_t c0 = seg_chunk0[sgi], c1 = seg_chunk0[sgi + 1];
    const uint32_t n = c1 - c0, per = (n + 63) / 64;
    const uint32_t a = c0 + min(n, l * per), b = c0 + min(n, (l + 1) * per);
    // (A)
    uint8_t* x = comp[l];
    uint8_t* y = work[l];
    uint32_t lx = 0;
    for (uint32_t c = b; c-- > a;) {
        const uint32_t ly = prepend_sig(sig + (size_t)c * 256, siglen[c], x, lx, y);
        uint8_t* t = x;
        x = y;
        y = t;
        lx = ly;
    }
    if (x != comp[l])
        for (uint32_t q = 0; q < lx; ++q) comp[l][q] = x[q];
    clen[l] = (uint16_t)lx;
    __syncthreads();
    // (B)
    if (l == 0) {
        ilen[63] = 0;
        for (int k = 62; k >= 0; --k)
            ilen[k] = (uint16_t)prepend_sig(comp[k + 1], clen[k + 1], inl[k + 1], ilen[k + 1], inl[k]);
        const uint32_t lt = prepend_sig(comp[0], clen[0], inl[0], ilen[0], work[0]);
        for (uint32_t q = 0; q < lt; ++q) table[(size_t)sgi * 256 + q] = work[0][q];
        nsym[sgi] = (uint16_t)lt;
    }
    __syncthreads();
    // (C)
    x = inl[l];
    uint8_t* y2 = comp[l];
    uint8_t* z = work[l];
    lx = ilen[l];
    for (uint32_t c = b; c-- > a;) {
        uint8_t* st = state + (size_t)c * 256;
        for (uint32_t q = 0; q < lx; ++q) st[q] = x[q];
        statelen[c] = (uint16_t)lx;
        if (c == a) break;
        const uint32_t ly = prepend_sig(sig + (size_t)c * 256, siglen[c], x, lx, y2);
        x = y2;
        lx = ly;
        uint8_t* t = y2;
        y2 = z;
        z = t;
    }
}

// One thread per chunk: replays its runs backwards from its starting list.
__global__ __launch_bounds__(kRankThreads) void k_chunk_rank(const uint8_t* __restrict__ rsym,
                                                            const Chunk* __restrict__ chunks, uint32_t nchunks,
                                                            const uint8_t* __restrict__ state,
                                                            const uint16_t* __restrict__ statelen,
                                                            uint8_t* __restrict__ rrank)
{
    __shared__ uint8_t lists[kRankThreads][256];
    const uint32_t c = blockIdx.x * kRankThreads + threadIdx.x;
    if (c >= nchunks) return;
    uint8_t* L = lists[threadIdx.x];
    const Chunk ch = chunks[c];
    uint32_t len = statelen[c];
    const uint8_t* st = state + (size_t)c * 256;
    for (uint32_t q = 0; q < len; ++q) L[q] = st[q];
    for (uint32_t r = ch.r1; r-- > ch.r0;) {
        const uint8_t s = rsym[r];
        uint32_t p = 0;
        while (p < len && L[p] != s) ++p;
        rrank[r] = (uint8_t)p;
        if (p == len) ++len;
        for (uint32_t q = p; q > 0; --q) L[q] = L[q - 1];
        L[0] = s;
    }
}

// ---------------------------------------------------------------------------
int QlfcFront::reserve(uint64_t n, uint32_t nblk)
{
    if (n <= cap_ && nblk <= cap_blk_) return 0;
    // grow geometrically so a stream of slowly growing batches reallocates
    // only a few times
    const uint64_t c = std::max<uint64_t>(std::max<uint64_t>(n, 1024), cap_ ? cap_ + cap_ / 2 : 0);
    const uint32_t cb = std::max<uint32_t>(std::max<uint32_t>(nblk, 1), cap_blk_ ? cap_blk_ + cap_blk_ / 2 : 0);
    release();
    const uint64_t nseg = (uint64_t)cb * 8;
    const uint64_t nchunks = c / kChunk + nseg + 1;
    QZ_HIP(hipMalloc(&flags_, c * 4));
    QZ_HIP(hipMalloc(&scan_, c * 4));
    QZ_HIP(hipMalloc(&tmp_, (scan_tmp_words(c) + 64) * 4));
    QZ_HIP(hipMalloc(&segflag_, c + 64));
    QZ_HIP(hipMalloc(&sblk_, cb * sizeof(SplitBlock)));
    QZ_HIP(hipMalloc(&segstart_, cb * 8 * 4));
    QZ_HIP(hipMalloc(&rstart_, c * 4));
    QZ_HIP(hipMalloc(&rsym_, c + 64));
    QZ_HIP(hipMalloc(&rrank_, c + 64));
    QZ_HIP(hipMalloc(&rlen_, c * 4));
    QZ_HIP(hipMalloc(&chunks_, nchunks * sizeof(Chunk)));
    QZ_HIP(hipMalloc(&segchunk_, (nseg + 1) * 4));
    QZ_HIP(hipMalloc(&sig_, nchunks * 256));
    QZ_HIP(hipMalloc(&siglen_, nchunks * 2));
    QZ_HIP(hipMalloc(&state_, nchunks * 256));
    QZ_HIP(hipMalloc(&statelen_, nchunks * 2));
    QZ_HIP(hipMalloc(&table_, nseg * 256));
    QZ_HIP(hipMalloc(&nsym_, nseg * 2));
    QZ_HIP(hipMalloc(&idx_, (nseg + 1) * 4));
    QZ_HIP(hipMalloc(&gath_, (nseg + 1) * 4));
    cap_ = c;
    cap_blk_ = cb;
    cap_chunks_ = nchunks;
    return 0;
}

void QlfcFront::release()
{
    void* ps[] = {flags_, scan_, tmp_, segflag_, sblk_, segstart_, rstart_, rsym_, rrank_, rlen_,
                  chunks_, segchunk_, sig_, siglen_, state_, statelen_, table_, nsym_, idx_, gath_};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    flags_ = scan_ = tmp_ = segstart_ = rstart_ = rlen_ = segchunk_ = idx_ = gath_ = nullptr;
    segflag_ = rsym_ = rrank_ = sig_ = state_ = table_ = nullptr;
    siglen_ = statelen_ = nsym_ = nullptr;
    sblk_ = chunks_ = nullptr;
    cap_ = 0;
    cap_blk_ = 0;
}

int QlfcFront::run(const uint8_t* U, const std::vector<uint32_t>& offs, hipStream_t s)
{
    const uint32_t nblk = (uint32_t)offs.size() - 1, N = offs.back();
    QZ_TRY(reserve(N, nblk));
    // 1. segment split
    std::vector<SplitBlock> sb(nblk);
    uint32_t nsamp = 0;
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t n = offs[b + 1] - offs[b];
        sb[b].off = offs[b];
        sb[b].n = n;
        sb[b].sbase = nsamp;
        sb[b].ns = n > 1 ? (n - 2) / 32 + 1 : 0;
        sb[b].nb = (uint32_t)num_segments_host((int)n);
        nsamp += sb[b].ns;
    }
    std::vector<uint32_t> segst((size_t)nblk * 8, 0xffffffffu);
    QZ_HIP(hipMemcpyAsync(segstart_, segst.data(), segst.size() * 4, hipMemcpyHostToDevice, s));
    if (nsamp) {
        QZ_HIP(hipMemcpyAsync(sblk_, sb.data(), nblk * sizeof(SplitBlock), hipMemcpyHostToDevice, s));
        const uint32_t g = qz_div_up(nsamp, kFThreads);
        hipLaunchKernelGGL(k_split_flags, dim3(g), dim3(kFThreads), 0, s, U, (const SplitBlock*)sblk_, nblk, nsamp,
                           flags_);
        QZ_HIP(hipGetLastError());
        QZ_TRY((device_scan<OpAdd, true>(flags_, scan_, nsamp, tmp_, s)));
        hipLaunchKernelGGL(k_split_find, dim3(g), dim3(kFThreads), 0, s, (const SplitBlock*)sblk_, nblk, nsamp,
                           flags_, scan_, segstart_);
        QZ_HIP(hipGetLastError());
        QZ_HIP(hipMemcpyAsync(segst.data(), segstart_, segst.size() * 4, hipMemcpyDeviceToHost, s));
        QZ_HIP(hipStreamSynchronize(s));
    }
    // segment list (global starts); a block whose boundaries were not all
    // found (changes <= nb) splits evenly
    segs_.clear();
    for (uint32_t b = 0; b < nblk; ++b) {
        const uint32_t n = sb[b].n, nb = sb[b].nb;
        bool found = true;
        for (uint32_t m = 1; m < nb; ++m) found = found && segst[b * 8 + m] != 0xffffffffu;
        for (uint32_t m = 0; m < nb; ++m) {
            uint32_t st, en;
            if (nb == 1) {
                st = 0;
                en = n;
            } else if (found) {
                st = m == 0 ? 0 : segst[b * 8 + m];
                en = m + 1 == nb ? n : segst[b * 8 + m + 1];
            } else {
                st = (n / nb) * m;
                en = m + 1 == nb ? n : (n / nb) * (m + 1);
            }
            segs_.push_back(Seg{b, offs[b] + st, en - st, 0, 0});
        }
    }
    // 2. run heads
    const uint32_t nseg = (uint32_t)segs_.size();
    std::vector<uint32_t> pos(nseg + 1);
    for (uint32_t i = 0; i < nseg; ++i) pos[i] = segs_[i].start;
    pos[nseg] = N - 1;
    QZ_HIP(hipMemsetAsync(segflag_, 0, N, s));
    QZ_HIP(hipMemcpyAsync(idx_, pos.data(), (nseg + 1) * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_mark, dim3(qz_div_up(nseg, 256)), dim3(256), 0, s, idx_, nseg, segflag_);
    QZ_HIP(hipGetLastError());
    const uint32_t g = qz_div_up(N, kFThreads);
    hipLaunchKernelGGL(k_run_heads, dim3(g), dim3(kFThreads), 0, s, U, N, segflag_, flags_);
    QZ_HIP(hipGetLastError());
    QZ_TRY((device_scan<OpAdd, true>(flags_, scan_, N, tmp_, s)));
    hipLaunchKernelGGL(k_run_emit, dim3(g), dim3(kFThreads), 0, s, U, N, flags_, scan_, rstart_, rsym_);
    QZ_HIP(hipGetLastError());
    // run index of every segment start (and the total)
    std::vector<uint32_t> rid(nseg + 1);
    hipLaunchKernelGGL(k_gather, dim3(qz_div_up(nseg + 1, 256)), dim3(256), 0, s, scan_, idx_, nseg + 1, gath_);
    QZ_HIP(hipGetLastError());
    QZ_HIP(hipMemcpyAsync(rid.data(), gath_, (nseg + 1) * 4, hipMemcpyDeviceToHost, s));
    QZ_HIP(hipStreamSynchronize(s));
    runs_ = rid[segs_.size()];
    for (size_t i = 0; i < segs_.size(); ++i) {
        segs_[i].r0 = rid[i] - 1;
        segs_[i].r1 = i + 1 < segs_.size() ? rid[i + 1] - 1 : runs_;
    }
    hipLaunchKernelGGL(k_run_len, dim3(qz_div_up(runs_, kFThreads)), dim3(kFThreads), 0, s, rstart_, runs_, N, rlen_);
    QZ_HIP(hipGetLastError());
    // 3. chunks
    std::vector<Chunk> ch;
    std::vector<uint32_t> sc(segs_.size() + 1);
    for (size_t i = 0; i < segs_.size(); ++i) {
        sc[i] = (uint32_t)ch.size();
        for (uint32_t r = segs_[i].r0; r < segs_[i].r1; r += kChunk) {
            const uint32_t e = std::min(r + kChunk, segs_[i].r1);
            ch.push_back(Chunk{r, e, (uint32_t)i, e == segs_[i].r1 ? 1u : 0u});
        }
    }
    sc[segs_.size()] = (uint32_t)ch.size();
    if (ch.size() > cap_chunks_) return QZ_BAD_PARAMETER;
    const uint32_t nch = (uint32_t)ch.size();
    QZ_HIP(hipMemcpyAsync(chunks_, ch.data(), nch * sizeof(Chunk), hipMemcpyHostToDevice, s));
    QZ_HIP(hipMemcpyAsync(segchunk_, sc.data(), sc.size() * 4, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_chunk_sig, dim3(qz_div_up(nch, kFThreads)), dim3(kFThreads), 0, s, rsym_,
                       (const Chunk*)chunks_, nch, sig_, siglen_);
    QZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_chunk_state, dim3(nseg), dim3(64), 0, s, segchunk_, sig_, siglen_, state_, statelen_,
                       table_, nsym_);
    QZ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_chunk_rank, dim3(qz_div_up(nch, kRankThreads)), dim3(kRankThreads), 0, s, rsym_,
                       (const Chunk*)chunks_, nch, state_, statelen_, rrank_);
    QZ_HIP(hipGetLastError());
    return 0;
}

}  // namespace qz
