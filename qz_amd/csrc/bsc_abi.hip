// libbsc drop-in: bsc_init / bsc_compress / bsc_block_info / bsc_decompress
// with the signatures, error codes and buffer contract qz-lib binds
// (crates/qz-lib/src/compression/bsc.rs:91-125).
//
// Block layout (libbsc's, pinned by the libbsc archive under tests/golden):
//   +0  i32 block size (header included)   +4  i32 n
//   +8  i32 mode = sorter | coder << 5 | lzpMinLen << 8 | lzpHashSize << 16
//       (coder: QZBSC_CODER_QZ_ADAPTIVE = 4 for every block this library
//       codes, since its rank/run model is not libbsc's QLFC; libbsc's own
//       ids 1-3 parse but their payloads are LIBBSC_NOT_SUPPORTED here)
//   +12 i32 primary index                  +16 adler32(data)
//   +20 adler32(bytes 28..end)             +24 adler32(bytes 0..23)
//   28.. coder payload: [nb u8][nb x (in i32, out i32)][segments]
//        (a single segment: [1][segment]), then num_indexes x i32 sampled
//        BWT indexes and the count byte num_indexes.
// Blocks of <= 28 bytes, or that do not shrink, are stored (mode 0).
// With LZP (mode bits 8..23 set) the data is LZP-encoded first (lzp.cpp),
// the coder payload is preceded by the LZP output size m (i32), and the
// BWT, primary and sampled indexes refer to those m bytes; a block that LZP
// does not shrink is coded without it. Requests for the static (1), adaptive
// (2) or fast (3) QLFC coder are all served by this library's adaptive model,
// and the header records coder 4, which is what the payload holds.
//
// The forward and inverse BWT always run on the GPU (there is no CPU
// transform in this library); the QLFC coder runs on the host.
#include "bsc_block.h"
#include "api_internal.h"
#include "engine.h"
#include "pool.h"
#include "lzp.h"
#include "qlfc_coder.h"
#include "../../include/qzbsc.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

namespace qz {

uint32_t adler32(const uint8_t* p, size_t n)
{
    uint32_t a = 1, b = 0;
    while (n > 0) {
        size_t k = n < 5552 ? n : 5552;
        n -= k;
        while (k >= 8) {
            a += p[0]; b += a; a += p[1]; b += a; a += p[2]; b += a; a += p[3]; b += a;
            a += p[4]; b += a; a += p[5]; b += a; a += p[6]; b += a; a += p[7]; b += a;
            p += 8;
            k -= 8;
        }
        while (k--) { a += *p++; b += a; }
        a %= 65521u;
        b %= 65521u;
    }
    return a | (b << 16);
}

static inline void put32(uint8_t* p, uint32_t v) { memcpy(p, &v, 4); }
static inline uint32_t get32(const uint8_t* p)
{
    uint32_t v;
    memcpy(&v, p, 4);
    return v;
}

int bwt_sample_rate(int n)
{
    int m = n / 8;
    m |= m >> 1; m |= m >> 2; m |= m >> 4; m |= m >> 8; m |= m >> 16;
    m >>= 1;
    return m + 1;
}

int store_block(const uint8_t* in, uint8_t* out, int n)
{
    const uint32_t a = adler32(in, (size_t)n);
    memmove(out + kHeader, in, (size_t)n);
    put32(out + 0, (uint32_t)(n + kHeader));
    put32(out + 4, (uint32_t)n);
    put32(out + 8, 0);
    put32(out + 12, 0);
    put32(out + 16, a);
    put32(out + 20, a);
    put32(out + 24, adler32(out, 24));
    return n + kHeader;
}

// Assembles the coder payload from per-segment coded bytes; a segment that
// did not code smaller than its input is stored raw (fetched via raw_seg).
template <class RawSeg>
static int assemble_payload(int n, int nb, const int* size, std::vector<std::vector<uint8_t>>& segout,
                            const std::vector<int>& res, RawSeg raw_seg, uint8_t* out, int cap)
{
    if (nb == 1) {
        if (res[0] < 0 || 1 + res[0] > cap) return QZ_NOT_COMPRESSIBLE;
        out[0] = 1;
        memcpy(out + 1, segout[0].data(), (size_t)res[0]);
        return 1 + res[0];
    }
    int ptr = 1 + 8 * nb;
    if (ptr > cap) return QZ_NOT_COMPRESSIBLE;
    out[0] = (uint8_t)nb;
    for (int s = 0; s < nb; ++s) {
        int r = res[s];
        const bool raw = r < 0 || r >= size[s];
        if (raw) r = size[s];
        if (ptr + r >= n || ptr + r > cap) return QZ_NOT_COMPRESSIBLE;
        put32(out + 1 + 8 * s, (uint32_t)size[s]);
        put32(out + 5 + 8 * s, (uint32_t)r);
        if (raw) {
            if (raw_seg(s, out + ptr)) return QZ_GPU_ERROR;
        } else {
            memcpy(out + ptr, segout[s].data(), (size_t)r);
        }
        ptr += r;
    }
    return ptr;
}

template <class Fn>
static void for_segments(int nb, bool parallel, Fn fn)
{
    if (parallel && nb > 1) {
        std::vector<std::thread> th;
        for (int s = 0; s < nb; ++s) th.emplace_back(fn, s);
        for (auto& t : th) t.join();
    } else {
        for (int s = 0; s < nb; ++s) fn(s);
    }
}

// Runs fn(b) for b in [0, n) on up to hardware_concurrency host threads.
template <class Fn>
static void parallel_blocks(int n, Fn fn)
{
    const int T = std::min<int>(n, (int)std::max(1u, std::thread::hardware_concurrency()));
    if (T <= 1) {
        for (int b = 0; b < n; ++b) fn(b);
        return;
    }
    std::atomic<int> next{0};
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&] {
            for (int b; (b = next.fetch_add(1)) < n;) fn(b);
        });
    for (auto& x : th) x.join();
}

// Coder payload for one block from its BWT (host front end). cap: bytes
// available.
int coder_compress(const uint8_t* bwt, int n, uint8_t* out, int cap, bool parallel)
{
    const int nb = qlfc::num_segments(n);
    int start[qlfc::kMaxSegments], size[qlfc::kMaxSegments];
    if (nb == 1) {
        start[0] = 0;
        size[0] = n;
    } else {
        qlfc::split_segments(bwt, n, nb, start, size);
    }
    std::vector<std::vector<uint8_t>> segout(nb);
    std::vector<int> res(nb, -1);
    auto work = [&](int s) {
        const int sn = size[s];
        std::vector<uint8_t> sym(sn), rank(sn);
        std::vector<uint32_t> len(sn);
        qlfc::SegRuns R;
        R.count = qlfc::front_end(bwt + start[s], sn, sym.data(), rank.data(), len.data(), R.table, &R.nsym);
        R.sym = sym.data();
        R.rank = rank.data();
        R.len = len.data();
        segout[s].resize((size_t)sn + 64);
        res[s] = qlfc::encode_segment(R, sn, segout[s].data(), sn);
    };
    for_segments(nb, parallel, work);
    return assemble_payload(
        n, nb, size, segout, res,
        [&](int sg, uint8_t* dst) {
            memcpy(dst, bwt + start[sg], (size_t)size[sg]);
            return 0;
        },
        out, cap);
}

// Host copy of the front end's results for a whole batch (after front.run()):
// the segment list, per run symbol, rank and length, per segment symbol table
// and symbol count. The runs also give the BWT bytes of a segment that is
// stored raw. It outlives the next front.run() on the slot.
// The runs (6 bytes each: length, symbol, rank) live in one pinned buffer
// from the cache, or in pageable memory if none can be pinned.
struct FrontHost {
    std::vector<QlfcFront::Seg> segs;
    std::vector<uint8_t> tables;
    std::vector<uint16_t> nsym;
    const uint32_t* len = nullptr;
    const uint8_t* sym = nullptr;
    const uint8_t* rank = nullptr;

    FrontHost() = default;
    FrontHost(const FrontHost&) = delete;
    FrontHost& operator=(const FrontHost&) = delete;
    FrontHost(FrontHost&& o) noexcept { *this = std::move(o); }
    FrontHost& operator=(FrontHost&& o) noexcept
    {
        if (this != &o) {
            free_runs();
            segs = std::move(o.segs);
            tables = std::move(o.tables);
            nsym = std::move(o.nsym);
            heap_ = std::move(o.heap_);
            pin_ = o.pin_;
            pin_cap_ = o.pin_cap_;
            len = o.len;
            sym = o.sym;
            rank = o.rank;
            o.pin_ = nullptr;
            o.pin_cap_ = 0;
            o.len = nullptr;
            o.sym = o.rank = nullptr;
        }
        return *this;
    }
    ~FrontHost() { free_runs(); }

    // Storage for R runs, laid out lengths | symbols | ranks.
    uint8_t* alloc_runs(size_t R)
    {
        free_runs();
        const size_t bytes = R * 6;
        uint8_t* p = bytes ? pinned_acquire(bytes, &pin_cap_) : nullptr;
        if (p) {
            pin_ = p;
        } else {
            heap_.resize(bytes / 4 + 2);
            p = reinterpret_cast<uint8_t*>(heap_.data());
        }
        len = reinterpret_cast<const uint32_t*>(p);
        sym = p + 4 * R;
        rank = p + 5 * R;
        return p;
    }

private:
    void free_runs()
    {
        if (pin_) pinned_release(pin_, pin_cap_);
        pin_ = nullptr;
        pin_cap_ = 0;
        std::vector<uint32_t>().swap(heap_);
        len = nullptr;
        sym = rank = nullptr;
    }
    uint8_t* pin_ = nullptr;
    size_t pin_cap_ = 0;
    std::vector<uint32_t> heap_;
};

static int front_to_host(const QlfcFront& front, FrontHost& fh, hipStream_t stream)
{
    const uint32_t R = front.runs_;  // runs of every segment, laid out in segment order
    const size_t S = front.segs_.size();
    fh.segs = front.segs_;
    uint8_t* p = fh.alloc_runs(R);
    fh.tables.resize(S * 256);
    fh.nsym.resize(S);
    if (R) {
        QZ_HIP(hipMemcpyAsync(p, front.rlen_, (size_t)R * 4, hipMemcpyDeviceToHost, stream));
        QZ_HIP(hipMemcpyAsync(p + (size_t)R * 4, front.rsym_, R, hipMemcpyDeviceToHost, stream));
        QZ_HIP(hipMemcpyAsync(p + (size_t)R * 5, front.rrank_, R, hipMemcpyDeviceToHost, stream));
    }
    if (S) {
        QZ_HIP(hipMemcpyAsync(fh.tables.data(), front.table_, S * 256, hipMemcpyDeviceToHost, stream));
        QZ_HIP(hipMemcpyAsync(fh.nsym.data(), front.nsym_, S * 2, hipMemcpyDeviceToHost, stream));
    }
    QZ_HIP(hipStreamSynchronize(stream));
    return 0;
}

// The BWT bytes of segment `seg`, rebuilt from its runs (a segment the
// coder cannot shrink is stored raw).
static int raw_from_runs(const FrontHost& fh, int seg, uint8_t* dst)
{
    const auto& sg = fh.segs[seg];
    uint64_t w = 0;
    for (uint32_t r = sg.r0; r < sg.r1; ++r) {
        const uint32_t L = fh.len[r];
        if (w + L > sg.size) return QZ_GPU_ERROR;
        memset(dst + w, fh.sym[r], L);
        w += L;
    }
    return w == sg.size ? 0 : QZ_GPU_ERROR;
}

// One coded segment of a batch.
struct SegCode {
    std::vector<uint8_t> out;
    int res = -1;
};

// Codes every segment of the batch from the host copy of the front end, all
// segments of all blocks in parallel.
static void code_segment(const FrontHost& fh, int i, SegCode& sc)
{
    const auto& sg = fh.segs[i];
    qlfc::SegRuns Rn;
    Rn.sym = fh.sym + sg.r0;
    Rn.rank = fh.rank + sg.r0;
    Rn.len = fh.len + sg.r0;
    Rn.count = sg.r1 - sg.r0;
    Rn.nsym = fh.nsym[i];
    memcpy(Rn.table, fh.tables.data() + (size_t)i * 256, 256);
    sc.out.resize((size_t)sg.size + 64);
    sc.res = qlfc::encode_segment(Rn, (int)sg.size, sc.out.data(), (int)sg.size);
}

// Coder payload of batch block `block` from its coded segments (each block's
// segments are taken once).
static int payload_from_segments(const FrontHost& fh, std::vector<SegCode>& sc,
                                 uint32_t block, int n, uint8_t* out, int cap)
{
    std::vector<int> sidx;
    for (int i = 0; i < (int)fh.segs.size(); ++i)
        if (fh.segs[i].block == block) sidx.push_back(i);
    const int nb = (int)sidx.size();
    if (nb < 1 || nb > qlfc::kMaxSegments) return QZ_GPU_ERROR;
    int size[qlfc::kMaxSegments], start[qlfc::kMaxSegments];
    std::vector<std::vector<uint8_t>> segout(nb);
    std::vector<int> res(nb);
    for (int k = 0; k < nb; ++k) {
        size[k] = (int)fh.segs[sidx[k]].size;
        start[k] = (int)fh.segs[sidx[k]].start;
        segout[k] = std::move(sc[sidx[k]].out);
        res[k] = sc[sidx[k]].res;
    }
    return assemble_payload(
        n, nb, size, segout, res,
        [&](int sg, uint8_t* dst) {
            return raw_from_runs(fh, sidx[sg], dst);
        },
        out, cap);
}

// Coder payload of batch block `block` from the host copy of the front end.
static int coder_compress_host(const FrontHost& fh, uint32_t block, int n, uint8_t* out,
                               int cap, bool parallel)
{
    std::vector<int> sidx;
    for (int i = 0; i < (int)fh.segs.size(); ++i)
        if (fh.segs[i].block == block) sidx.push_back(i);
    const int nb = (int)sidx.size();
    if (nb < 1 || nb > qlfc::kMaxSegments) return QZ_GPU_ERROR;
    int size[qlfc::kMaxSegments], start[qlfc::kMaxSegments];
    for (int k = 0; k < nb; ++k) {
        size[k] = (int)fh.segs[sidx[k]].size;
        start[k] = (int)fh.segs[sidx[k]].start;
    }
    std::vector<std::vector<uint8_t>> segout(nb);
    std::vector<int> res(nb, -1);
    for_segments(nb, parallel, [&](int k) {
        const auto& sg = fh.segs[sidx[k]];
        qlfc::SegRuns Rn;
        Rn.sym = fh.sym + sg.r0;
        Rn.rank = fh.rank + sg.r0;
        Rn.len = fh.len + sg.r0;
        Rn.count = sg.r1 - sg.r0;
        Rn.nsym = fh.nsym[sidx[k]];
        memcpy(Rn.table, fh.tables.data() + (size_t)sidx[k] * 256, 256);
        segout[k].resize((size_t)sg.size + 64);
        res[k] = qlfc::encode_segment(Rn, (int)sg.size, segout[k].data(), (int)sg.size);
    });
    return assemble_payload(
        n, nb, size, segout, res,
        [&](int sg, uint8_t* dst) {
            return raw_from_runs(fh, sidx[sg], dst);
        },
        out, cap);
}

// One segment of a coder payload to decode: in[0..isz) -> out[0..osz). A
// segment of a multi-segment payload whose coded size equals its size is
// stored raw.
struct SegTask {
    const uint8_t* in;
    int isz;
    uint8_t* out;
    int osz;
    bool raw_ok;
};

// Parses the segment table of a coder payload that decodes into out (n
// bytes) and appends its segments to t.
static int coder_tasks(const uint8_t* in, int in_size, uint8_t* out, int n, std::vector<SegTask>& t)
{
    if (in_size < 1) return QZ_UNEXPECTED_EOB;
    const int nb = in[0];
    if (nb < 1 || nb > qlfc::kMaxSegments) return QZ_DATA_CORRUPT;
    if (nb == 1) {
        t.push_back({in + 1, in_size - 1, out, n, false});
        return 0;
    }
    if (1 + 8 * nb > in_size) return QZ_UNEXPECTED_EOB;
    int64_t ip = 1 + 8 * nb, op = 0;
    SegTask seg[qlfc::kMaxSegments];
    for (int s = 0; s < nb; ++s) {
        const int osz = (int)get32(in + 1 + 8 * s), isz = (int)get32(in + 5 + 8 * s);
        if (osz < 1 || isz < 1) return QZ_DATA_CORRUPT;
        if (ip + isz > in_size || op + osz > n) return QZ_DATA_CORRUPT;
        seg[s] = {in + ip, isz, out + op, osz, true};
        ip += isz;
        op += osz;
    }
    if (op != n) return QZ_DATA_CORRUPT;
    t.insert(t.end(), seg, seg + nb);
    return 0;
}

static int run_seg_task(const SegTask& x)
{
    if (x.raw_ok && x.isz == x.osz) {
        memcpy(x.out, x.in, (size_t)x.isz);
        return 0;
    }
    return qlfc::decode_segment(x.in, x.isz, x.out, x.osz) ? QZ_DATA_CORRUPT : 0;
}

int coder_decompress(const uint8_t* in, int in_size, uint8_t* out, int n, bool parallel)
{
    std::vector<SegTask> t;
    QZ_TRY(coder_tasks(in, in_size, out, n, t));
    std::vector<int> rc(t.size(), 0);
    for_segments((int)t.size(), parallel, [&](int s) { rc[s] = run_seg_task(t[s]); });
    for (int r : rc)
        if (r) return QZ_DATA_CORRUPT;
    return 0;
}

int check_params(int n, int lzpHashSize, int lzpMinLen, int blockSorter, int coder, int* mode)
{
    if (blockSorter != LIBBSC_BLOCKSORTER_BWT) return (blockSorter >= 3 && blockSorter <= 8) ? QZ_NOT_SUPPORTED
                                                                                              : QZ_BAD_PARAMETER;
    if (coder < 1 || coder > 3) return QZ_BAD_PARAMETER;
    if ((lzpMinLen != 0 || lzpHashSize != 0) && !lzp::params_ok(lzpHashSize, lzpMinLen)) return QZ_BAD_PARAMETER;
    if (n < 0 || n > (1 << 30)) return QZ_BAD_PARAMETER;
    *mode = blockSorter | (QZBSC_CODER_QZ_ADAPTIVE << 5);
    if (lzpHashSize) *mode |= (lzpMinLen << 8) | (lzpHashSize << 16);
    return 0;
}

// Finishes a block whose BWT (in `bwt`, n bytes), primary index and sampled
// indexes are known: codes the payload into out+28 and writes the header.
int finish_block(const uint8_t* input, const uint8_t* bwt, int n, int mode, int primary, const int* aux,
                 uint8_t* out, bool parallel)
{
    const uint32_t adler_data = adler32(input, (size_t)n);
    int num_indexes = (n - 1) / bwt_sample_rate(n);
    if (n < 64 * 1024) num_indexes = 0;
    std::vector<uint8_t> payload((size_t)n + 4096);
    const int result = coder_compress(bwt, n, payload.data(), (int)payload.size(), parallel);
    if (result < 0 || result + 1 + 4 * num_indexes >= n) return store_block(input, out, n);
    memcpy(out + kHeader, payload.data(), (size_t)result);
    for (int t = 0; t < num_indexes; ++t) put32(out + kHeader + result + 4 * t, (uint32_t)(aux[t + 1] - 1));
    out[kHeader + result + 4 * num_indexes] = (uint8_t)num_indexes;
    const int body = result + 1 + 4 * num_indexes;
    put32(out + 0, (uint32_t)(body + kHeader));
    put32(out + 4, (uint32_t)n);
    put32(out + 8, (uint32_t)mode);
    put32(out + 12, (uint32_t)primary);
    put32(out + 16, adler_data);
    put32(out + 20, adler32(out + kHeader, (size_t)body));
    put32(out + 24, adler32(out, 24));
    return body + kHeader;
}

// Same as finish_block, with the coder payload produced by a callable. dn
// is the size of the data the BWT ran on: n, or the LZP output size when
// mode carries LZP parameters (then written ahead of the coder payload).
template <class Payload>
static int finish_block_with(const uint8_t* input, int n, int dn, int mode, int primary, const int* aux,
                             uint8_t* out, Payload payload_fn)
{
    const uint32_t adler_data = adler32(input, (size_t)n);
    int num_indexes = (dn - 1) / bwt_sample_rate(dn);
    if (dn < 64 * 1024) num_indexes = 0;
    const int pre = (mode >> 8) ? 4 : 0;
    std::vector<uint8_t> payload((size_t)dn + 4096);
    const int result = payload_fn(payload.data(), (int)payload.size());
    if (result == QZ_GPU_ERROR) return result;
    if (result < 0 || pre + result + 1 + 4 * num_indexes >= n) return store_block(input, out, n);
    if (pre) put32(out + kHeader, (uint32_t)dn);
    memcpy(out + kHeader + pre, payload.data(), (size_t)result);
    for (int t = 0; t < num_indexes; ++t) put32(out + kHeader + pre + result + 4 * t, (uint32_t)(aux[t + 1] - 1));
    out[kHeader + pre + result + 4 * num_indexes] = (uint8_t)num_indexes;
    const int body = pre + result + 1 + 4 * num_indexes;
    put32(out + 0, (uint32_t)(body + kHeader));
    put32(out + 4, (uint32_t)n);
    put32(out + 8, (uint32_t)mode);
    put32(out + 12, (uint32_t)primary);
    put32(out + 16, adler_data);
    put32(out + 20, adler32(out + kHeader, (size_t)body));
    put32(out + 24, adler32(out, 24));
    return body + kHeader;
}

int parse_header(const uint8_t* h, int size, BlockHeader* bh)
{
    if (size < kHeader) return QZ_UNEXPECTED_EOB;
    if (adler32(h, 24) != get32(h + 24)) return QZ_DATA_CORRUPT;
    bh->block_size = (int)get32(h);
    bh->data_size = (int)get32(h + 4);
    bh->mode = (int)get32(h + 8);
    bh->index = (int)get32(h + 12);
    bh->adler_data = get32(h + 16);
    bh->adler_body = get32(h + 20);
    const int sorter = bh->mode & 0x1f, coder = (bh->mode >> 5) & 0x7;
    const int lzpMin = (bh->mode >> 8) & 0xff, lzpHash = (bh->mode >> 16) & 0xff;
    bool ok = bh->block_size >= kHeader && bh->data_size >= 0 && bh->index >= 0 && bh->index <= bh->data_size;
    if (bh->mode != 0) {
        ok = ok && sorter >= 1 && sorter <= 8 && coder >= 1 && coder <= QZBSC_CODER_QZ_ADAPTIVE &&
             (bh->mode >> 24) == 0;
        ok = ok && ((lzpMin == 0 && lzpHash == 0) || (lzpMin >= 4 && lzpHash >= 10 && lzpHash <= 28));
    } else {
        ok = ok && bh->block_size == bh->data_size + kHeader;
    }
    return ok ? 0 : QZ_DATA_CORRUPT;
}

// Checks the body checksum and the mode of a block and appends its coder
// segments (decoding into `bwt`, n = data size) to t. For a stored block
// copies the data and returns 1.
static int payload_tasks(const uint8_t* input, int inputSize, const BlockHeader& bh, uint8_t* bwt,
                         std::vector<SegTask>& t)
{
    if (inputSize < bh.block_size) return QZ_UNEXPECTED_EOB;
    if (adler32(input + kHeader, (size_t)(bh.block_size - kHeader)) != bh.adler_body) return QZ_DATA_CORRUPT;
    if (bh.mode == 0) {
        memmove(bwt, input + kHeader, (size_t)bh.data_size);
        return 1;
    }
    const int sorter = bh.mode & 0x1f, coder = (bh.mode >> 5) & 0x7;
    if (sorter != LIBBSC_BLOCKSORTER_BWT || coder != QZBSC_CODER_QZ_ADAPTIVE || (bh.mode >> 8) != 0)
        return QZ_NOT_SUPPORTED;
    if (bh.data_size < 1 || bh.index < 1) return QZ_DATA_CORRUPT;
    const int num_indexes = input[bh.block_size - 1];
    const int payload = bh.block_size - kHeader - 1 - 4 * num_indexes;
    if (payload < 1) return QZ_DATA_CORRUPT;
    return coder_tasks(input + kHeader, payload, bwt, bh.data_size, t);
}

// Checks the body checksum and decodes the coder payload into `bwt`
// (n = data size). For a stored block copies the data and returns 1.
int decode_payload(const uint8_t* input, int inputSize, const BlockHeader& bh, uint8_t* bwt, bool parallel)
{
    std::vector<SegTask> t;
    const int r = payload_tasks(input, inputSize, bh, bwt, t);
    if (r != 0) return r;
    std::vector<int> rc(t.size(), 0);
    for_segments((int)t.size(), parallel, [&](int k) { rc[k] = run_seg_task(t[k]); });
    for (int x : rc)
        if (x) return QZ_DATA_CORRUPT;
    return 0;
}

// Device batches of the batch entry points hold at most this many bytes.
constexpr uint64_t kBatchBytes = 1ull << 30;
// Batch pipelines: device sub-batches of about 1/pipe_parts() of a batch's
// bytes, none smaller than kPipeBytes. QZBSC_PIPE_PARTS (1-64) overrides the
// default count.
constexpr uint64_t kPipeParts = 4, kPipeBytes = 32ull << 20;
static uint64_t pipe_parts()
{
    const char* e = getenv("QZBSC_PIPE_PARTS");
    const long v = e ? strtol(e, nullptr, 10) : 0;
    return (v >= 1 && v <= 64) ? (uint64_t)v : kPipeParts;
}

// LZP blocks: the size m of the LZP data the BWT ran on (0 if the block
// carries no LZP parameters), checked against the header.
int lzp_data_size(const uint8_t* input, const BlockHeader& bh)
{
    if ((bh.mode >> 8) == 0 || bh.mode == 0) return 0;
    if ((bh.mode & 0x1f) != LIBBSC_BLOCKSORTER_BWT || ((bh.mode >> 5) & 0x7) != QZBSC_CODER_QZ_ADAPTIVE)
        return QZ_NOT_SUPPORTED;  // libbsc's LZP + QLFC formats are not this library's
    if (bh.block_size < kHeader + 4) return QZ_DATA_CORRUPT;
    const int m = (int)get32(input + kHeader);
    if (m < 1 || m >= bh.data_size || bh.index > m) return QZ_DATA_CORRUPT;
    return m;
}

}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API int bsc_init(int features)
{
    (void)features;
    qlfc::num_segments(1);
    return pool_device_ok() ? LIBBSC_NO_ERROR : QZBSC_GPU_ERROR;
}

QZBSC_API int bsc_compress(const unsigned char* input, unsigned char* output, int n, int lzpHashSize, int lzpMinLen,
                           int blockSorter, int coder, int features)
{
    int mode = 0;
    QZ_TRY(check_params(n, lzpHashSize, lzpMinLen, blockSorter, coder, &mode));
    if (!input || !output) return QZ_BAD_PARAMETER;
    if (n <= kHeader) return store_block(input, output, n);
    // LZP first (host); a block it does not shrink is coded without it
    std::vector<uint8_t> lz;
    const uint8_t* data = input;
    int dn = n;
    if (mode >> 8) {
        lz.resize((size_t)n);
        const int m = lzp::encode(input, n, lz.data(), n, lzpHashSize, lzpMinLen);
        if (m > 1) {
            data = lz.data();
            dn = m;
        } else {
            mode &= 0xff;
        }
    }
    std::vector<uint32_t> offs = {0u, (uint32_t)dn};
    int primary = 0;
    std::vector<int> aux(256);
    FrontHost fh;
    {
        // the device slot is held for the device work only: the host coder
        // below runs with it released, so concurrent callers keep the device
        // busy while this one codes
        SlotGuard g;
        if (!g.ok()) return QZ_GPU_ERROR;
        Slot& sl = g.slot();
        QZ_TRY(sl.fwd.reserve((uint64_t)dn, 1));
        QZ_HIP(hipMemcpyAsync(sl.fwd.text_, data, (size_t)dn, hipMemcpyHostToDevice, sl.stream));
        QZ_TRY(fwd_batch_device(sl, sl.fwd.text_, sl.fwd.out_, offs, &primary, aux.data(), nullptr));
        QZ_TRY(sl.front.run(sl.fwd.out_, offs, sl.stream));
        QZ_TRY(front_to_host(sl.front, fh, sl.stream));
    }
    const bool par = (features & LIBBSC_FEATURE_MULTITHREADING) != 0;
    return finish_block_with(input, n, dn, mode, primary, aux.data(), output, [&](uint8_t* p, int cap) {
        return coder_compress_host(fh, 0, dn, p, cap, par);
    });
}

QZBSC_API int bsc_block_info(const unsigned char* blockHeader, int headerSize, int* pBlockSize, int* pDataSize,
                             int features)
{
    (void)features;
    if (!blockHeader) return QZ_BAD_PARAMETER;
    BlockHeader bh;
    QZ_TRY(parse_header(blockHeader, headerSize, &bh));
    if (pBlockSize) *pBlockSize = bh.block_size;
    if (pDataSize) *pDataSize = bh.data_size;
    return LIBBSC_NO_ERROR;
}

QZBSC_API int bsc_decompress(const unsigned char* input, int inputSize, unsigned char* output, int outputSize,
                             int features)
{
    if (!input || !output) return QZ_BAD_PARAMETER;
    BlockHeader bh;
    QZ_TRY(parse_header(input, inputSize, &bh));
    if (inputSize < bh.block_size || outputSize < bh.data_size) return QZ_UNEXPECTED_EOB;
    const bool par = (features & LIBBSC_FEATURE_MULTITHREADING) != 0;
    const int m = lzp_data_size(input, bh);
    if (m < 0) return m;
    if (m > 0) {
        // LZP block: coder payload after the size word -> m-byte BWT ->
        // inverse BWT -> LZP decode into the caller's buffer
        if (adler32(input + kHeader, (size_t)(bh.block_size - kHeader)) != bh.adler_body) return QZ_DATA_CORRUPT;
        const int num_indexes = input[bh.block_size - 1];
        const int payload = bh.block_size - kHeader - 4 - 1 - 4 * num_indexes;
        if (payload < 1) return QZ_DATA_CORRUPT;
        std::vector<uint8_t> tmp((size_t)m);
        if (coder_decompress(input + kHeader + 4, payload, tmp.data(), m, par)) return QZ_DATA_CORRUPT;
        if (m > 1) QZ_TRY(qzbsc_bwt_decode_batch(tmp.data(), tmp.data(), &m, 1, &bh.index));
        const int lzpMin = (bh.mode >> 8) & 0xff, lzpHash = (bh.mode >> 16) & 0xff;
        if (lzp::decode(tmp.data(), m, output, bh.data_size, lzpHash, lzpMin)) return QZ_DATA_CORRUPT;
    } else {
        const int rc = decode_payload(input, inputSize, bh, output, par);
        if (rc < 0) return rc;
        if (rc == 0 && bh.data_size > 1) {
            int n = bh.data_size;
            QZ_TRY(qzbsc_bwt_decode_batch(output, output, &n, 1, &bh.index));
        }
    }
    if (adler32(output, (size_t)bh.data_size) != bh.adler_data) return QZ_DATA_CORRUPT;
    return LIBBSC_NO_ERROR;
}

// Batch form of bsc_compress. Block b (sizes[b] bytes) starts at input +
// sum of the earlier sizes; its output goes to output + sum over c < b of
// (sizes[c] + 28 + 1024), libbsc's per-block buffer contract, and its size
// (or a negative libbsc code) to out_sizes[b]. Device sub-batches (up to
// kBatchBytes each) sort the suffixes of their blocks and run the front end;
// the host codes the segments of one sub-batch in parallel while the device
// runs the next. Returns 0, or the first block's error.
QZBSC_API int qzbsc_compress_batch(const unsigned char* input, const int* sizes, int nblocks, unsigned char* output,
                                   int* out_sizes, int lzpHashSize, int lzpMinLen, int blockSorter, int coder,
                                   int features)
{
    if (!input || !sizes || !output || !out_sizes || nblocks < 1) return QZ_BAD_PARAMETER;
    std::vector<int> mode(nblocks), dn(nblocks);
    std::vector<size_t> ioff(nblocks + 1, 0), ooff(nblocks + 1, 0);
    for (int b = 0; b < nblocks; ++b) {
        QZ_TRY(check_params(sizes[b], lzpHashSize, lzpMinLen, blockSorter, coder, &mode[b]));
        ioff[b + 1] = ioff[b] + (size_t)sizes[b];
        ooff[b + 1] = ooff[b] + (size_t)sizes[b] + kHeader + 1024;
    }
    (void)features;
    // LZP first (host, per block); a block it does not shrink is coded without it
    std::vector<std::vector<uint8_t>> lz(nblocks);
    parallel_blocks(nblocks, [&](int b) {
        const int n = sizes[b];
        dn[b] = n;
        if (n > kHeader && (mode[b] >> 8)) {
            lz[b].resize((size_t)n);
            const int m = lzp::encode(input + ioff[b], n, lz[b].data(), n, lzpHashSize, lzpMinLen);
            if (m > 1) {
                dn[b] = m;
            } else {
                lz[b].clear();
                mode[b] &= 0xff;
            }
        }
    });
    bool device = false;
    for (int b = 0; b < nblocks; ++b) {
        if (sizes[b] <= kHeader) out_sizes[b] = store_block(input + ioff[b], output + ooff[b], sizes[b]);
        else device = true;
    }
    if (!device) return LIBBSC_NO_ERROR;  // only stored blocks: no device needed
    std::unique_ptr<SlotGuard> guard(new SlotGuard);
    if (!guard->ok()) return QZ_GPU_ERROR;
    Slot& sl = guard->slot();
    // Device sub-batches of about 1/pipe_parts() of the bytes (at least
    // kPipeBytes, at most kBatchBytes). After each one this thread queues its
    // coder segments, longest first, for a pool of coder threads that runs
    // for the whole call; the thread that codes a block's last segment
    // assembles that block, and the one that assembles a sub-batch's last
    // block frees its host copy of the front end.
    struct Job {
        std::vector<int> list, prim, aux;
        FrontHost fh;
        std::vector<SegCode> sc;
        std::vector<int> left;  // uncoded segments per block of the sub-batch
        int blocks_left = 0;
    };
    // QZBSC_TRACE_BATCH=1 prints, per sub-batch, the time of its upload,
    // device work and front-end download, and when the coder pool finished.
    const bool trace = getenv("QZBSC_TRACE_BATCH") != nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    auto run_device = [&](Job& job) -> int {
        const std::vector<int>& list = job.list;
        const int k = (int)list.size();
        std::vector<uint32_t> offs(k + 1, 0);
        for (int i = 0; i < k; ++i) offs[i + 1] = offs[i] + (uint32_t)dn[list[i]];
        QZ_TRY(sl.fwd.reserve(offs[k], (uint32_t)k, BwtEngine::max_block(offs)));
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < k; ++i) {
            const int b = list[i];
            const uint8_t* data = lz[b].empty() ? input + ioff[b] : lz[b].data();
            QZ_HIP(hipMemcpyAsync(sl.fwd.text_ + offs[i], data, (size_t)dn[b], hipMemcpyHostToDevice, sl.stream));
        }
        if (trace) QZ_HIP(hipStreamSynchronize(sl.stream));
        const double up = ms_since(t0);
        job.prim.assign(k, 0);
        job.aux.assign((size_t)k * 256, 0);
        QZ_TRY(fwd_batch_device(sl, sl.fwd.text_, sl.fwd.out_, offs, job.prim.data(), job.aux.data(), nullptr));
        if (trace) QZ_HIP(hipStreamSynchronize(sl.stream));
        const double sort = ms_since(t0) - up;
        QZ_TRY(sl.front.run(sl.fwd.out_, offs, sl.stream));
        if (trace) QZ_HIP(hipStreamSynchronize(sl.stream));
        const double front = ms_since(t0) - up - sort;
        const int rc = front_to_host(sl.front, job.fh, sl.stream);
        if (trace)
            fprintf(stderr, "[qzbsc batch] sub-batch of %d blocks, %u B: upload %.1f ms, sort %.1f ms, front end %.1f ms, "
                            "runs %u down %.1f ms (at %.1f ms)\n",
                    k, offs[k], up, sort, front, sl.front.runs_, ms_since(t0) - up - sort - front, ms_since(t_start));
        return rc;
    };
    std::vector<Job> jobs;
    {
        uint64_t total = 0;
        for (int b = 0; b < nblocks; ++b)
            if (sizes[b] > kHeader) total += (uint64_t)dn[b];
        const uint64_t target = std::min<uint64_t>(kBatchBytes, std::max<uint64_t>(kPipeBytes, total / pipe_parts()));
        uint64_t bytes = 0;
        for (int b = 0; b < nblocks; ++b) {
            if (sizes[b] <= kHeader) continue;
            if (jobs.empty() || (bytes + (uint64_t)dn[b] > target && !jobs.back().list.empty())) {
                jobs.emplace_back();
                bytes = 0;
            }
            jobs.back().list.push_back(b);
            bytes += (uint64_t)dn[b];
        }
    }
    auto finish = [&](Job& job, int i) {
        const int b = job.list[i];
        out_sizes[b] = finish_block_with(input + ioff[b], sizes[b], dn[b], mode[b], job.prim[i],
                                         job.aux.data() + (size_t)i * 256, output + ooff[b],
                                         [&](uint8_t* p, int cap) {
                                             return payload_from_segments(job.fh, job.sc, (uint32_t)i, dn[b], p, cap);
                                         });
    };
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<int, int>> queue;  // (sub-batch, segment)
    bool closed = false;
    auto worker = [&] {
        for (;;) {
            std::pair<int, int> t;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv.wait(lk, [&] { return !queue.empty() || closed; });
                if (queue.empty()) return;
                t = queue.front();
                queue.pop_front();
            }
            Job& job = jobs[t.first];
            code_segment(job.fh, t.second, job.sc[t.second]);
            const int i = (int)job.fh.segs[t.second].block;
            bool block_done, job_done = false;
            {
                std::lock_guard<std::mutex> lk(mu);
                block_done = --job.left[i] == 0;
            }
            if (!block_done) continue;
            finish(job, i);
            {
                std::lock_guard<std::mutex> lk(mu);
                job_done = --job.blocks_left == 0;
            }
            if (job_done) {
                job.fh = FrontHost{};
                job.sc = std::vector<SegCode>{};
            }
        }
    };
    std::vector<std::thread> pool;
    int nsegs = 0;  // coder segments of the whole batch (the front end's rule)
    for (int b = 0; b < nblocks; ++b)
        if (sizes[b] > kHeader) nsegs += qlfc::num_segments(dn[b]);
    const int T = std::min<int>(nsegs, (int)std::max(1u, std::thread::hardware_concurrency()));
    for (int t = 0; t < std::max(T, 1); ++t) pool.emplace_back(worker);
    int dev_rc = 0;
    for (size_t j = 0; j < jobs.size() && !dev_rc; ++j) {
        Job& job = jobs[j];
        dev_rc = run_device(job);
        if (dev_rc) break;
        const int S = (int)job.fh.segs.size(), k = (int)job.list.size();
        job.sc.assign(S, SegCode{});
        job.left.assign(k, 0);
        for (int i = 0; i < S; ++i) ++job.left[job.fh.segs[i].block];
        job.blocks_left = k;
        std::vector<int> order(S);
        for (int i = 0; i < S; ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return job.fh.segs[x].size > job.fh.segs[y].size; });
        // blocks without segments (the front end gives every block one) are
        // finished here, before any worker can see this sub-batch
        for (int i = 0; i < k; ++i)
            if (!job.left[i]) {
                finish(job, i);
                --job.blocks_left;
            }
        if (!S) continue;
        {
            std::lock_guard<std::mutex> lk(mu);
            for (int i : order) queue.push_back({(int)j, i});
        }
        cv.notify_all();
    }
    guard.reset();  // the device work is done: other callers may take the slot
    {
        std::lock_guard<std::mutex> lk(mu);
        closed = true;
    }
    cv.notify_all();
    for (auto& t : pool) t.join();
    if (trace) fprintf(stderr, "[qzbsc batch] coder pool done at %.1f ms\n", ms_since(t_start));
    if (dev_rc) return dev_rc;
    for (int b = 0; b < nblocks; ++b)
        if (out_sizes[b] < 0) return out_sizes[b];
    return LIBBSC_NO_ERROR;
}

// Batch form of bsc_decompress. Block b (in_sizes[b] bytes) starts at input
// + sum of the earlier in_sizes and decodes into output + sum of the earlier
// out_sizes, out_sizes[b] being at least its data size (bsc_block_info).
// The host decodes the coder segments in parallel, sub-batch by sub-batch,
// and each device sub-batch (up to kBatchBytes) inverts its BWTs as soon as
// its segments are decoded; each block's result (0 or a negative libbsc
// code) goes to rcs[b]. Returns 0, or the first block's error.
QZBSC_API int qzbsc_decompress_batch(const unsigned char* input, const int* in_sizes, int nblocks,
                                     unsigned char* output, const int* out_sizes, int features, int* rcs)
{
    if (!input || !in_sizes || !output || !out_sizes || !rcs || nblocks < 1) return QZ_BAD_PARAMETER;
    (void)features;
    std::vector<size_t> ioff(nblocks + 1, 0), ooff(nblocks + 1, 0);
    for (int b = 0; b < nblocks; ++b) {
        if (in_sizes[b] < 0 || out_sizes[b] < 0) return QZ_BAD_PARAMETER;
        ioff[b + 1] = ioff[b] + (size_t)in_sizes[b];
        ooff[b + 1] = ooff[b] + (size_t)out_sizes[b];
    }
    std::vector<BlockHeader> bh(nblocks);
    std::vector<std::vector<uint8_t>> tmp(nblocks);  // LZP blocks: the m-byte BWT domain
    std::vector<int> need(nblocks, 0);               // BWT bytes to invert
    std::vector<uint8_t*> bw(nblocks, nullptr);
    std::vector<std::vector<SegTask>> tasks(nblocks);  // coder segments per block
    parallel_blocks(nblocks, [&](int b) {
        const uint8_t* in = input + ioff[b];
        uint8_t* out = output + ooff[b];
        int rc = parse_header(in, in_sizes[b], &bh[b]);
        if (!rc && (in_sizes[b] < bh[b].block_size || out_sizes[b] < bh[b].data_size)) rc = QZ_UNEXPECTED_EOB;
        const int m = rc ? 0 : lzp_data_size(in, bh[b]);
        if (!rc && m < 0) rc = m;
        if (!rc && m > 0) {
            if (adler32(in + kHeader, (size_t)(bh[b].block_size - kHeader)) != bh[b].adler_body) {
                rc = QZ_DATA_CORRUPT;
            } else {
                const int num_indexes = in[bh[b].block_size - 1];
                const int payload = bh[b].block_size - kHeader - 4 - 1 - 4 * num_indexes;
                tmp[b].resize((size_t)m);
                if (payload < 1 || coder_tasks(in + kHeader + 4, payload, tmp[b].data(), m, tasks[b]))
                    rc = QZ_DATA_CORRUPT;
                else if (m > 1) {
                    bw[b] = tmp[b].data();
                    need[b] = m;
                }
            }
        } else if (!rc) {
            const int r = payload_tasks(in, in_sizes[b], bh[b], out, tasks[b]);
            if (r < 0) rc = r;
            else if (r == 0 && bh[b].data_size > 1) {
                bw[b] = out;
                need[b] = bh[b].data_size;
            }
        }
        rcs[b] = rc;
    });
    // Device sub-batches: the blocks to invert, in order, each sub-batch
    // about 1/kPipeParts of their BWT bytes (at least kPipeBytes, at most
    // kBatchBytes). Segment tasks run in sub-batch order, longest first within
    // one, on a host pool while this thread inverts each sub-batch on the
    // device as soon as its last segment is decoded.
    std::vector<int> stage(nblocks, 0);
    std::vector<std::vector<int>> batches;
    {
        uint64_t total = 0;
        for (int b = 0; b < nblocks; ++b) {
            if (rcs[b] || !need[b]) continue;
            if (bh[b].index < 1 || bh[b].index > need[b]) {  // a primary the inverse cannot take
                rcs[b] = QZ_DATA_CORRUPT;
                continue;
            }
            total += (uint64_t)need[b];
        }
        const uint64_t target = std::min<uint64_t>(kBatchBytes, std::max<uint64_t>(kPipeBytes, total / pipe_parts()));
        uint64_t bytes = 0;
        for (int b = 0; b < nblocks; ++b) {
            if (rcs[b] || !need[b]) continue;
            if (batches.empty() || (bytes + (uint64_t)need[b] > target && !batches.back().empty())) {
                batches.emplace_back();
                bytes = 0;
            }
            batches.back().push_back(b);
            bytes += (uint64_t)need[b];
            stage[b] = (int)batches.size() - 1;
        }
    }
    // QZBSC_TRACE_BATCH=1 prints, per sub-batch, when its segments were
    // decoded and the time of its upload, inverse and download.
    const bool trace = getenv("QZBSC_TRACE_BATCH") != nullptr;
    const auto t_start = std::chrono::steady_clock::now();
    auto ms_since = [](std::chrono::steady_clock::time_point t) {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    };
    auto flush = [&](Slot& sl, const std::vector<int>& list) -> int {
        const int k = (int)list.size();
        std::vector<uint32_t> offs(k + 1, 0);
        std::vector<int> prim(k);
        for (int i = 0; i < k; ++i) {
            offs[i + 1] = offs[i] + (uint32_t)need[list[i]];
            prim[i] = bh[list[i]].index;
        }
        QZ_TRY(sl.inv.reserve(offs[k], (uint32_t)k));
        const double t_dec = ms_since(t_start);
        const auto t0 = std::chrono::steady_clock::now();
        for (int i = 0; i < k; ++i)
            QZ_HIP(hipMemcpyAsync(sl.inv.U_ + offs[i], bw[list[i]], (size_t)need[list[i]], hipMemcpyHostToDevice,
                                  sl.stream));
        if (trace) QZ_HIP(hipStreamSynchronize(sl.stream));
        const double up = ms_since(t0);
        QZ_TRY(inv_batch_device(sl, sl.inv.U_, sl.inv.T_, offs, prim.data(), nullptr));
        if (trace) QZ_HIP(hipStreamSynchronize(sl.stream));
        const double inv = ms_since(t0) - up;
        for (int i = 0; i < k; ++i)
            QZ_HIP(hipMemcpyAsync(bw[list[i]], sl.inv.T_ + offs[i], (size_t)need[list[i]], hipMemcpyDeviceToHost,
                                  sl.stream));
        QZ_HIP(hipStreamSynchronize(sl.stream));
        if (trace)
            fprintf(stderr, "[qzbsc batch] decoded %d blocks, %u B by %.1f ms: upload %.1f ms, inverse %.1f ms, "
                            "download %.1f ms\n",
                    k, offs[k], t_dec, up, inv, ms_since(t0) - up - inv);
        return 0;
    };
    std::unique_ptr<SlotGuard> guard;
    if (!batches.empty()) {  // stored and single-byte blocks need no device
        guard.reset(new SlotGuard);
        if (!guard->ok()) return QZ_GPU_ERROR;
    }
    struct Item {
        int b;
        const SegTask* t;
    };
    std::vector<Item> all;
    for (int b = 0; b < nblocks; ++b)
        for (const auto& t : tasks[b]) all.push_back({b, &t});
    std::stable_sort(all.begin(), all.end(), [&](const Item& x, const Item& y) {
        return stage[x.b] != stage[y.b] ? stage[x.b] < stage[y.b] : x.t->osz > y.t->osz;
    });
    std::vector<int> left(std::max<size_t>(1, batches.size()), 0);  // undecoded segments per sub-batch
    for (const auto& it : all) ++left[stage[it.b]];
    std::vector<uint8_t> bad(nblocks, 0);
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<int> next{0};
    std::vector<std::thread> pool;
    const int T = std::min<int>((int)all.size(), (int)std::max(1u, std::thread::hardware_concurrency()));
    for (int t = 0; t < T; ++t)
        pool.emplace_back([&] {
            for (int i; (i = next.fetch_add(1)) < (int)all.size();) {
                const int r = run_seg_task(*all[i].t);
                std::lock_guard<std::mutex> lk(mu);
                if (r) bad[all[i].b] = 1;
                if (--left[stage[all[i].b]] == 0) cv.notify_all();
            }
        });
    int dev_rc = 0;
    size_t failed_from = batches.size();  // first sub-batch the device did not invert
    for (size_t j = 0; j < batches.size(); ++j) {
        std::vector<int> list;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return left[j] == 0; });
            for (int b : batches[j])
                if (!bad[b]) list.push_back(b);
        }
        if (!list.empty() && !dev_rc) {
            dev_rc = flush(guard->slot(), list);
            if (dev_rc) failed_from = j;
        }
    }
    for (auto& t : pool) t.join();
    for (int b = 0; b < nblocks; ++b)
        if (bad[b] && !rcs[b]) rcs[b] = QZ_DATA_CORRUPT;
    if (dev_rc) {
        // the failed sub-batch and every later one were never inverted: their
        // output still holds BWT bytes, so each of their blocks reports the error
        for (size_t j = failed_from; j < batches.size(); ++j)
            for (int b : batches[j])
                if (!rcs[b]) rcs[b] = dev_rc;
        return dev_rc;
    }
    parallel_blocks(nblocks, [&](int b) {
        if (rcs[b]) return;
        uint8_t* out = output + ooff[b];
        if (!tmp[b].empty()) {
            const int lzpMin = (bh[b].mode >> 8) & 0xff, lzpHash = (bh[b].mode >> 16) & 0xff;
            if (lzp::decode(tmp[b].data(), (int)tmp[b].size(), out, bh[b].data_size, lzpHash, lzpMin)) {
                rcs[b] = QZ_DATA_CORRUPT;
                return;
            }
        }
        if (adler32(out, (size_t)bh[b].data_size) != bh[b].adler_data) rcs[b] = QZ_DATA_CORRUPT;
    });
    if (trace) fprintf(stderr, "[qzbsc batch] checks done at %.1f ms\n", ms_since(t_start));
    for (int b = 0; b < nblocks; ++b)
        if (rcs[b]) return rcs[b];
    return LIBBSC_NO_ERROR;
}

// LZP stage alone (host): libbsc lzp.h's bsc_lzp_compress / _decompress.
QZBSC_API int qzbsc_lzp_compress(const unsigned char* input, unsigned char* output, int n, int lzpHashSize,
                                 int lzpMinLen)
{
    return lzp::encode(input, n, output, n, lzpHashSize, lzpMinLen);
}

QZBSC_API int qzbsc_lzp_decompress(const unsigned char* input, int inputSize, unsigned char* output, int outputSize,
                                   int lzpHashSize, int lzpMinLen)
{
    return lzp::decode(input, inputSize, output, outputSize, lzpHashSize, lzpMinLen);
}

// Device QLFC front end of one BWT block (host buffers): segments, tables
// and per-run symbol / rank / length. Returns the run count.
QZBSC_API int qzbsc_qlfc_front(const unsigned char* bwt, int n, int* nseg, int* seg_start, int* seg_size,
                               int* seg_runs, unsigned char* tables, int* nsym, unsigned char* run_sym,
                               unsigned char* run_rank, int* run_len)
{
    if (!bwt || n < 1 || !nseg || !seg_start || !seg_size || !seg_runs || !tables || !nsym || !run_sym ||
        !run_rank || !run_len)
        return QZ_BAD_PARAMETER;
    SlotGuard g;
    if (!g.ok()) return QZ_GPU_ERROR;
    Slot& sl = g.slot();
    QZ_TRY(sl.fwd.reserve((uint64_t)n, 1));
    QZ_HIP(hipMemcpyAsync(sl.fwd.out_, bwt, (size_t)n, hipMemcpyHostToDevice, sl.stream));
    std::vector<uint32_t> offs = {0u, (uint32_t)n};
    QZ_TRY(sl.front.run(sl.fwd.out_, offs, sl.stream));
    const auto& segs = sl.front.segs_;
    const uint32_t R = sl.front.runs_;
    *nseg = (int)segs.size();
    std::vector<uint16_t> ns(segs.size());
    for (size_t i = 0; i < segs.size(); ++i) {
        seg_start[i] = (int)segs[i].start;
        seg_size[i] = (int)segs[i].size;
        seg_runs[i] = (int)(segs[i].r1 - segs[i].r0);
    }
    std::vector<uint32_t> len(R);
    QZ_HIP(hipMemcpyAsync(tables, sl.front.table_, segs.size() * 256, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipMemcpyAsync(ns.data(), sl.front.nsym_, segs.size() * 2, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipMemcpyAsync(run_sym, sl.front.rsym_, R, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipMemcpyAsync(run_rank, sl.front.rrank_, R, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipMemcpyAsync(len.data(), sl.front.rlen_, (size_t)R * 4, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipStreamSynchronize(sl.stream));
    for (size_t i = 0; i < segs.size(); ++i) nsym[i] = ns[i];
    for (uint32_t r = 0; r < R; ++r) run_len[r] = (int)len[r];
    return (int)R;
}

// Test hook: size word and symbol table of one coded segment.
QZBSC_API int qzbsc_segment_header(const unsigned char* seg, int size, int* n, unsigned char* table, int* nsym)
{
    if (!seg || !n || !table || !nsym) return QZ_BAD_PARAMETER;
    return qlfc::read_segment_header(seg, size, n, table, nsym);
}

// Host-only coder round trip of one BWT-domain block (no device needed).
QZBSC_API int qzbsc_coder_compress(const unsigned char* bwt, int n, unsigned char* out, int cap, int features)
{
    if (!bwt || !out || n < 1) return QZ_BAD_PARAMETER;
    return coder_compress(bwt, n, out, cap, (features & LIBBSC_FEATURE_MULTITHREADING) != 0);
}

QZBSC_API int qzbsc_coder_decompress(const unsigned char* in, int size, unsigned char* out, int n, int features)
{
    if (!in || !out || n < 1) return QZ_BAD_PARAMETER;
    return coder_decompress(in, size, out, n, (features & LIBBSC_FEATURE_MULTITHREADING) != 0);
}

}  // extern "C"
