// C-ABI for the BWT stages: per-block calls with libbsc's bwt.h signatures,
// batch calls over host buffers, and device-resident batch calls for
// callers whose blocks already sit in HBM.
//
// Engines (each with its own stream and resident workspace) live in a pool,
// so concurrent callers (qz drives libbsc from many rayon workers,
// crates/qz-lib/src/compression/mod.rs:402-404) each get one.
#include "engine.h"
#include "pool.h"
#include "api_internal.h"
#include "../../include/qzbsc.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace qz {

static int sample_rate(int n)
{
    int m = n / 8;
    m |= m >> 1; m |= m >> 2; m |= m >> 4; m |= m >> 8; m |= m >> 16;
    m >>= 1;
    return m + 1;
}

// A batch of the BWT entries holds < kBatchMax bytes (32-bit offsets); the
// forward sorts it in chunks and the inverse in sub-batches of whole blocks
// below kSubBatchMax (its row entries carry a 31-bit LF), so blocks are
// below kSubBatchMax too.
constexpr uint64_t kBatchMax = 0xffff0000ull;
constexpr uint64_t kSubBatchMax = 0x7fff0000ull;

static int check_sizes(const int* sizes, int nblocks, std::vector<uint32_t>& offs)
{
    if (nblocks <= 0 || !sizes) return QZ_BAD_PARAMETER;
    offs.assign(nblocks + 1, 0);
    uint64_t tot = 0;
    for (int b = 0; b < nblocks; ++b) {
        if (sizes[b] < 1 || (uint64_t)sizes[b] >= kSubBatchMax) return QZ_BAD_PARAMETER;
        offs[b] = (uint32_t)tot;
        tot += (uint64_t)sizes[b];
        if (tot >= kBatchMax) return QZ_BAD_PARAMETER;
    }
    offs[nblocks] = (uint32_t)tot;
    return 0;
}

// Runs of consecutive blocks of offs holding < kSubBatchMax bytes each, as
// (first block, end block).
static std::vector<std::pair<uint32_t, uint32_t>> sub_batches(const std::vector<uint32_t>& offs)
{
    std::vector<std::pair<uint32_t, uint32_t>> parts;
    const uint32_t nblk = (uint32_t)offs.size() - 1;
    for (uint32_t b = 0; b < nblk;) {
        uint32_t e = b + 1;
        while (e < nblk && (uint64_t)offs[e + 1] - offs[b] < kSubBatchMax) ++e;
        parts.emplace_back(b, e);
        b = e;
    }
    return parts;
}

// Forward BWT of a batch whose text already sits in device memory.
int fwd_batch_device(Slot& sl, const uint8_t* d_in, uint8_t* d_out, const std::vector<uint32_t>& offs,
                            int* primary, int* aux, float* gpu_ms)
{
    const uint32_t nblk = (uint32_t)offs.size() - 1, n = offs.back();
    hipStream_t s = sl.stream;
    QZ_TRY(sl.fwd.reserve(n, nblk, BwtEngine::max_block(offs)));
    QZ_HIP(hipEventRecord(sl.ev0, s));
    std::vector<int32_t> idx;
    QZ_TRY(sl.fwd.encode_batch(d_in, d_out, offs, s, idx));
    QZ_HIP(hipEventRecord(sl.ev1, s));
    QZ_HIP(hipStreamSynchronize(s));
    if (gpu_ms) QZ_HIP(hipEventElapsedTime(gpu_ms, sl.ev0, sl.ev1));
    for (uint32_t b = 0; b < nblk; ++b) {
        const int nb = (int)(offs[b + 1] - offs[b]);
        const int cnt = (nb - 1) / sample_rate(nb);
        primary[b] = idx[(size_t)b * 256];
        if (aux)
            for (int k = 0; k <= cnt && k < 256; ++k) aux[(size_t)b * 256 + k] = idx[(size_t)b * 256 + k];
    }
    return 0;
}

int inv_batch_device(Slot& sl, const uint8_t* d_in, uint8_t* d_out, const std::vector<uint32_t>& offs,
                            const int* primary, float* gpu_ms)
{
    const uint32_t nblk = (uint32_t)offs.size() - 1, n = offs.back();
    for (uint32_t b = 0; b < nblk; ++b) {
        const int nb = (int)(offs[b + 1] - offs[b]);
        if (primary[b] < 1 || primary[b] > nb) return QZ_BAD_PARAMETER;
    }
    hipStream_t s = sl.stream;
    if (n < kSubBatchMax) {
        QZ_TRY(sl.inv.reserve(n, nblk));
        QZ_HIP(hipMemcpyAsync(sl.inv.off_, offs.data(), offs.size() * 4, hipMemcpyHostToDevice, s));
        QZ_HIP(hipMemcpyAsync(sl.inv.prim_, primary, (size_t)nblk * 4, hipMemcpyHostToDevice, s));
        QZ_HIP(hipEventRecord(sl.ev0, s));
        QZ_TRY(sl.inv.decode_device(d_in, d_out, n, nblk, offs, s));
    } else {
        // sub-batches of whole blocks, each with its own offsets; the
        // workspace is reserved for the largest
        const auto parts = sub_batches(offs);
        uint64_t mn = 0;
        uint32_t mb = 0;
        for (const auto& p : parts) {
            mn = std::max<uint64_t>(mn, offs[p.second] - offs[p.first]);
            mb = std::max(mb, p.second - p.first);
        }
        QZ_TRY(sl.inv.reserve(mn, mb));
        QZ_HIP(hipEventRecord(sl.ev0, s));
        std::vector<uint32_t> lo;
        for (const auto& p : parts) {
            const uint32_t base = offs[p.first];
            lo.assign(offs.begin() + p.first, offs.begin() + p.second + 1);
            for (uint32_t& o : lo) o -= base;
            QZ_HIP(hipMemcpyAsync(sl.inv.off_, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, s));
            QZ_HIP(hipMemcpyAsync(sl.inv.prim_, primary + p.first, (size_t)(p.second - p.first) * 4,
                                  hipMemcpyHostToDevice, s));
            QZ_TRY(sl.inv.decode_device(d_in + base, d_out + base, lo.back(), p.second - p.first, lo, s));
            QZ_HIP(hipStreamSynchronize(s));  // lo is rewritten for the next sub-batch
        }
    }
    QZ_HIP(hipEventRecord(sl.ev1, s));
    QZ_HIP(hipStreamSynchronize(s));
    if (gpu_ms) QZ_HIP(hipEventElapsedTime(gpu_ms, sl.ev0, sl.ev1));
    return 0;
}

}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API int qzbsc_bwt_encode_batch(const unsigned char* in, unsigned char* out, const int* sizes, int nblocks,
                                     int* primary, int* aux)
{
    std::vector<uint32_t> offs;
    QZ_TRY(check_sizes(sizes, nblocks, offs));
    if (!in || !out || !primary) return QZ_BAD_PARAMETER;
    SlotGuard g;
    if (!g.ok()) return QZ_GPU_ERROR;
    Slot& sl = g.slot();
    const uint32_t n = offs.back();
    QZ_TRY(sl.fwd.reserve(n, (uint32_t)nblocks, BwtEngine::max_block(offs)));
    QZ_HIP(hipMemcpyAsync(sl.fwd.text_, in, n, hipMemcpyHostToDevice, sl.stream));
    QZ_TRY(fwd_batch_device(sl, sl.fwd.text_, sl.fwd.out_, offs, primary, aux, nullptr));
    QZ_HIP(hipMemcpyAsync(out, sl.fwd.out_, n, hipMemcpyDeviceToHost, sl.stream));
    QZ_HIP(hipStreamSynchronize(sl.stream));
    return 0;
}

QZBSC_API int qzbsc_bwt_decode_batch(const unsigned char* in, unsigned char* out, const int* sizes, int nblocks,
                                     const int* primary)
{
    std::vector<uint32_t> offs;
    QZ_TRY(check_sizes(sizes, nblocks, offs));
    if (!in || !out || !primary) return QZ_BAD_PARAMETER;
    SlotGuard g;
    if (!g.ok()) return QZ_GPU_ERROR;
    Slot& sl = g.slot();
    // through the engine's staging one sub-batch of whole blocks at a time
    for (const auto& p : sub_batches(offs)) {
        const uint32_t base = offs[p.first], n = offs[p.second] - base;
        std::vector<uint32_t> lo(offs.begin() + p.first, offs.begin() + p.second + 1);
        for (uint32_t& o : lo) o -= base;
        QZ_TRY(sl.inv.reserve(n, p.second - p.first));
        QZ_HIP(hipMemcpyAsync(sl.inv.U_, in + base, n, hipMemcpyHostToDevice, sl.stream));
        QZ_TRY(inv_batch_device(sl, sl.inv.U_, sl.inv.T_, lo, primary + p.first, nullptr));
        QZ_HIP(hipMemcpyAsync(out + base, sl.inv.T_, n, hipMemcpyDeviceToHost, sl.stream));
        QZ_HIP(hipStreamSynchronize(sl.stream));
    }
    return 0;
}

QZBSC_API int qzbsc_dev_bwt_encode_batch(const void* d_in, void* d_out, const int* sizes, int nblocks, int* primary,
                                         int* aux, float* gpu_ms)
{
    std::vector<uint32_t> offs;
    QZ_TRY(check_sizes(sizes, nblocks, offs));
    if (!d_in || !d_out || !primary) return QZ_BAD_PARAMETER;
    SlotGuard g;
    if (!g.ok()) return QZ_GPU_ERROR;
    return fwd_batch_device(g.slot(), (const uint8_t*)d_in, (uint8_t*)d_out, offs, primary, aux, gpu_ms);
}

QZBSC_API int qzbsc_dev_bwt_decode_batch(const void* d_in, void* d_out, const int* sizes, int nblocks,
                                         const int* primary, float* gpu_ms)
{
    std::vector<uint32_t> offs;
    QZ_TRY(check_sizes(sizes, nblocks, offs));
    if (!d_in || !d_out || !primary) return QZ_BAD_PARAMETER;
    SlotGuard g;
    if (!g.ok()) return QZ_GPU_ERROR;
    return inv_batch_device(g.slot(), (const uint8_t*)d_in, (uint8_t*)d_out, offs, primary, gpu_ms);
}

// libbsc bwt.h: int bsc_bwt_encode(unsigned char* T, int n, unsigned char*
// num_indexes, int* indexes, int features). In place; returns the primary
// index; writes (n-1)/r trailer indexes I[k+1] - 1.
QZBSC_API int qzbsc_bwt_encode(unsigned char* T, int n, unsigned char* num_indexes, int* indexes, int features)
{
    (void)features;
    if (!T || n < 1 || !num_indexes || !indexes) return QZ_BAD_PARAMETER;
    int primary = 0;
    std::vector<int> aux(256);
    QZ_TRY(qzbsc_bwt_encode_batch(T, T, &n, 1, &primary, aux.data()));
    const int cnt = (n - 1) / sample_rate(n);
    *num_indexes = (unsigned char)cnt;
    for (int t = 0; t < cnt; ++t) indexes[t] = aux[t + 1] - 1;
    return primary;
}

// libbsc bwt.h: int bsc_bwt_decode(unsigned char* T, int n, int index,
// unsigned char num_indexes, int* indexes, int features). In place.
QZBSC_API int qzbsc_bwt_decode(unsigned char* T, int n, int index, unsigned char num_indexes, int* indexes,
                               int features)
{
    (void)num_indexes;
    (void)indexes;
    (void)features;
    if (!T || n < 0 || index <= 0 || index > n) return QZ_BAD_PARAMETER;
    if (n <= 1) return QZ_NO_ERROR;
    return qzbsc_bwt_decode_batch(T, T, &n, 1, &index);
}

// Digit-width tuner state (which = 0: round-0 sort, 1: large-group sort):
// chosen width (0 while undecided) and mean ns per element for 8 / 9 / 11 bits.
// Returns the number of timed trials.
QZBSC_API int qzbsc_tuning(int which, int* choice, double* ns_per_elem)
{
    if ((which != 0 && which != 1) || !choice || !ns_per_elem) return QZ_BAD_PARAMETER;
    return tune_state(which, choice, ns_per_elem);
}

QZBSC_API int qzbsc_tune_knob(int knob, int* choice, int* ncand, int* cand, double* ns_per_unit)
{
    if (knob < 0 || knob > 3) return -1;
    if (!choice || !ncand || !cand || !ns_per_unit) return QZ_BAD_PARAMETER;
    return tune_knob(knob, choice, ncand, cand, ns_per_unit);
}

QZBSC_API int qzbsc_graph_stats(unsigned long long* captures, unsigned long long* replays)
{
    if (!captures || !replays) return QZ_BAD_PARAMETER;
    *captures = g_graph_captures.load();
    *replays = g_graph_replays.load();
    return 0;
}

}  // extern "C"
