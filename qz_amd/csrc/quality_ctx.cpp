// quality_ctx: qz's context-adaptive quality coder, restated for the host
// (crates/qz-lib/src/compression/quality_ctx.rs). It keeps the reference's
// exact integer arithmetic, so blobs are interchangeable with qz's.
//
//   - range coder with a 64-bit low, cache and carry propagation, 2^24
//     normalisation, five flushing shifts (quality_ctx.rs:29-163);
//   - one adaptive frequency model per context, Laplace-initialised,
//     +1 per coded symbol, rescaled when its total reaches 2^20
//     (quality_ctx.rs:171-247). The rescale halves each frequency computed
//     against the already-updated lower bound, as the reference does;
//   - context = (position / 5 capped at 63, previous quality capped at 49,
//     |q[-1] - q[-2]| <= 2, previous base, current base), 160000 contexts
//     mapped to models in first-use order through 16-bit slot ids
//     (quality_ctx.rs:253-268, 347-356; a slot id is stored truncated to 16
//     bits, as there);
//   - blob: [n_symbols u8][symbols][read_len u16 LE, 0 = variable]
//     [num_reads u32 LE][compressed_len u32 LE][compressed]
//     (quality_ctx.rs:375-385).
#include "quality_ctx.h"

#include "common.h"
#include "../../include/qzbsc.h"

#include <cstring>
#include <vector>

namespace qz {
namespace qctx {

namespace {

constexpr uint32_t kTop = 1u << 24;
constexpr int kPosBin = 5, kMaxPosBins = 64, kMaxQual = 50, kBaseCtx = 25, kDelta = 2;
constexpr int kContexts = kMaxPosBins * kMaxQual * kDelta * kBaseCtx;  // 160000
constexpr uint32_t kRescale = 1u << 20;
constexpr uint16_t kNoSlot = 0xffff;

struct Encoder {
    uint64_t low = 0;
    uint32_t range = 0xffffffffu;
    uint8_t cache = 0;
    uint32_t cache_size = 1;
    std::vector<uint8_t> out;

    void shift_low()
    {
        const uint8_t hi = (uint8_t)(low >> 32);
        if (hi != 0 || (uint32_t)low < 0xff000000u) {
            uint8_t byte = cache;
            for (;;) {
                out.push_back((uint8_t)(byte + hi));
                byte = 0xff;
                if (--cache_size == 0) break;
            }
            cache = (uint8_t)((low >> 24) & 0xff);
        }
        ++cache_size;
        low = (uint64_t)((uint32_t)low << 8);
    }
    void encode(uint32_t cum, uint32_t freq, uint32_t total)
    {
        const uint32_t r = range / total;
        low += (uint64_t)cum * r;
        if (cum + freq < total) range = r * freq;
        else range -= r * cum;
        while (range < kTop) {
            range <<= 8;
            shift_low();
        }
    }
    void finish()
    {
        for (int i = 0; i < 5; ++i) shift_low();
    }
};

struct Decoder {
    uint32_t range = 0xffffffffu, code = 0;
    const uint8_t* in;
    size_t n, pos = 0;
    Decoder(const uint8_t* p, size_t len) : in(p), n(len)
    {
        if (n) pos = 1;  // the encoder's first byte is its initial cache
        for (int i = 0; i < 4; ++i) code = (code << 8) | next();
    }
    uint32_t next() { return pos < n ? in[pos++] : 0u; }
    int decode(const uint32_t* cum, int nsym, uint32_t total)
    {
        const uint32_t r = range / total;
        uint32_t off = code / r;
        if (off > total - 1) off = total - 1;
        int s = 0;
        while (s + 1 < nsym && cum[s + 1] <= off) ++s;
        const uint32_t c = cum[s], f = cum[s + 1] - c;
        code -= c * r;
        if (c + f < total) range = r * f;
        else range -= r * c;
        while (range < kTop) {
            code = (code << 8) | next();
            range <<= 8;
        }
        return s;
    }
};

struct Models {
    std::vector<uint32_t> cum;     // per model: nsym + 1 cumulative counts
    std::vector<uint32_t> totals;
    int nsym;
    explicit Models(int k) : nsym(k) {}
    size_t alloc()
    {
        const size_t slot = totals.size();
        for (int i = 0; i <= nsym; ++i) cum.push_back((uint32_t)i);
        totals.push_back((uint32_t)nsym);
        return slot;
    }
    uint32_t* at(size_t slot) { return cum.data() + slot * (size_t)(nsym + 1); }
    void update(size_t slot, int sym)
    {
        uint32_t* c = at(slot);
        for (int i = sym + 1; i <= nsym; ++i) ++c[i];
        if (++totals[slot] >= kRescale) rescale(slot);
    }
    void rescale(size_t slot)
    {
        uint32_t* c = at(slot);
        uint32_t acc = 0;
        c[0] = 0;
        for (int i = 0; i < nsym; ++i) {
            const uint32_t f = c[i + 1] - c[i];  // c[i] already holds the new bound
            const uint32_t nf = (f >> 1) > 1u ? (f >> 1) : 1u;
            acc += nf;
            c[i + 1] = acc;
        }
        totals[slot] = acc;
    }
};

inline uint8_t base_index(uint8_t b)
{
    switch (b) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return 4;
    }
}

inline int context_id(int pos, uint8_t pq1, uint8_t pq2, uint8_t pb, uint8_t cb)
{
    const int bin = pos / kPosBin < kMaxPosBins - 1 ? pos / kPosBin : kMaxPosBins - 1;
    const int pq = pq1 < kMaxQual - 1 ? pq1 : kMaxQual - 1;
    const int d = (int)pq1 - (int)pq2;
    const int db = (d >= -2 && d <= 2) ? 0 : 1;
    int bc = (pb < 4 ? pb : 4) * 5 + (cb < 4 ? cb : 4);
    if (bc > kBaseCtx - 1) bc = kBaseCtx - 1;
    return bin * (kMaxQual * kDelta * kBaseCtx) + pq * (kDelta * kBaseCtx) + db * kBaseCtx + bc;
}

// Model slot of a context, allocating on first use; ids are kept as 16 bits.
inline size_t slot_of(std::vector<uint16_t>& slots, Models& m, int ctx)
{
    if (slots[ctx] == kNoSlot) {
        const size_t id = m.alloc();
        slots[ctx] = (uint16_t)id;
        return id;
    }
    return slots[ctx];
}

inline void put16(std::vector<uint8_t>& v, uint32_t x)
{
    v.push_back((uint8_t)x);
    v.push_back((uint8_t)(x >> 8));
}
inline void put32v(std::vector<uint8_t>& v, uint32_t x)
{
    for (int i = 0; i < 4; ++i) v.push_back((uint8_t)(x >> (8 * i)));
}
inline uint32_t get_le(const uint8_t* p, int k)
{
    uint32_t x = 0;
    for (int i = 0; i < k; ++i) x |= (uint32_t)p[i] << (8 * i);
    return x;
}

}  // namespace

int compress(const uint8_t* quals, const int* qlen, const uint8_t* seqs, const int* slen, int nreads,
             std::vector<uint8_t>& blob)
{
    blob.clear();
    if (nreads == 0) return 0;
    if (nreads < 0 || !quals || !qlen || !seqs || !slen) return QZ_BAD_PARAMETER;
    bool seen[256] = {false};
    size_t qo = 0;
    for (int r = 0; r < nreads; ++r) {
        if (qlen[r] < 0 || slen[r] < 0) return QZ_BAD_PARAMETER;
        for (int j = 0; j < qlen[r]; ++j) {
            const uint8_t b = quals[qo + j];
            if (b < 33) return QZ_BAD_PARAMETER;
            seen[b - 33] = true;
        }
        qo += (size_t)qlen[r];
    }
    uint8_t symbols[256], sym_map[256] = {0};
    int nsym = 0;
    for (int i = 0; i < 256; ++i)
        if (seen[i]) {
            sym_map[i] = (uint8_t)nsym;
            symbols[nsym++] = (uint8_t)i;
        }
    if (nsym == 0) return QZ_BAD_PARAMETER;
    bool variable = false;
    for (int r = 1; r < nreads; ++r) variable |= qlen[r] != qlen[0];
    const uint32_t read_len = variable ? 0u : (uint32_t)(uint16_t)qlen[0];

    std::vector<uint16_t> slots(kContexts, kNoSlot);
    Models m(nsym);
    Encoder enc;
    size_t so = 0;
    qo = 0;
    for (int r = 0; r < nreads; ++r) {
        if (qlen[r] > slen[r]) return QZ_BAD_PARAMETER;
        uint8_t pq1 = 0, pq2 = 0, pb = 4;
        for (int j = 0; j < qlen[r]; ++j) {
            const uint8_t phred = (uint8_t)(quals[qo + j] - 33);
            const int sym = sym_map[phred];
            const uint8_t cb = base_index(seqs[so + j]);
            const size_t slot = slot_of(slots, m, context_id(j, pq1, pq2, pb, cb));
            const uint32_t* c = m.at(slot);
            enc.encode(c[sym], c[sym + 1] - c[sym], m.totals[slot]);
            m.update(slot, sym);
            pq2 = pq1;
            pq1 = phred;
            pb = cb;
        }
        qo += (size_t)qlen[r];
        so += (size_t)slen[r];
    }
    enc.finish();
    blob.push_back((uint8_t)nsym);
    blob.insert(blob.end(), symbols, symbols + nsym);
    put16(blob, read_len);
    put32v(blob, (uint32_t)nreads);
    put32v(blob, (uint32_t)enc.out.size());
    blob.insert(blob.end(), enc.out.begin(), enc.out.end());
    return 0;
}

int decompress(const uint8_t* data, size_t size, const uint8_t* seqs, const int* slen, int nreads, uint8_t* quals,
               int* qlen)
{
    if (nreads == 0) return 0;
    if (nreads < 0 || !seqs || !slen || !quals || !qlen) return QZ_BAD_PARAMETER;
    if (!data || size == 0) return QZ_DATA_CORRUPT;
    size_t pos = 0;
    auto need = [&](size_t k) { return pos + k <= size; };
    if (!need(1)) return QZ_DATA_CORRUPT;
    const int nsym = data[pos++];
    if (nsym == 0 || !need((size_t)nsym + 10)) return QZ_DATA_CORRUPT;
    const uint8_t* symbols = data + pos;
    pos += (size_t)nsym;
    const uint32_t read_len = get_le(data + pos, 2);
    pos += 2;
    if (get_le(data + pos, 4) != (uint32_t)nreads) return QZ_DATA_CORRUPT;
    pos += 4;
    const uint32_t clen = get_le(data + pos, 4);
    pos += 4;
    if (!need(clen)) return QZ_DATA_CORRUPT;
    Decoder dec(data + pos, clen);
    std::vector<uint16_t> slots(kContexts, kNoSlot);
    Models m(nsym);
    size_t so = 0, qo = 0;
    for (int r = 0; r < nreads; ++r) {
        const int len = read_len ? (int)read_len : slen[r];
        if (slen[r] < len) return QZ_DATA_CORRUPT;
        uint8_t pq1 = 0, pq2 = 0, pb = 4;
        for (int j = 0; j < len; ++j) {
            const uint8_t cb = base_index(seqs[so + j]);
            const size_t slot = slot_of(slots, m, context_id(j, pq1, pq2, pb, cb));
            const int sym = dec.decode(m.at(slot), nsym, m.totals[slot]);
            m.update(slot, sym);
            const uint8_t phred = symbols[sym];
            quals[qo + j] = (uint8_t)(phred + 33);
            pq2 = pq1;
            pq1 = phred;
            pb = cb;
        }
        qlen[r] = len;
        qo += (size_t)len;
        so += (size_t)slen[r];
    }
    return 0;
}

}  // namespace qctx
}  // namespace qz

using namespace qz;

extern "C" {

QZBSC_API long long qzbsc_quality_ctx_compress(const unsigned char* quals, const int* qual_lens,
                                               const unsigned char* seqs, const int* seq_lens, int num_reads,
                                               unsigned char* out, long long cap)
{
    std::vector<uint8_t> blob;
    const int rc = qctx::compress(quals, qual_lens, seqs, seq_lens, num_reads, blob);
    if (rc) return rc;
    if ((long long)blob.size() > cap || (!out && !blob.empty())) return QZ_NOT_ENOUGH_MEMORY;
    if (!blob.empty()) memcpy(out, blob.data(), blob.size());
    return (long long)blob.size();
}

QZBSC_API int qzbsc_quality_ctx_decompress(const unsigned char* data, long long size, const unsigned char* seqs,
                                           const int* seq_lens, int num_reads, unsigned char* quals, int* qual_lens)
{
    if (size < 0) return QZ_BAD_PARAMETER;
    return qctx::decompress(data, (size_t)size, seqs, seq_lens, num_reads, quals, qual_lens);
}

}  // extern "C"
