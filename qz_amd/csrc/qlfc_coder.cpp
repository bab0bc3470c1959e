// QLFC coder, host side (exact integer arithmetic).
//
// Segment bitstream (one adaptive model per segment, reset at its start):
//   32-bit size word | symbol table | per run: rank, then run length.
// The size word, the table coding (each table symbol sent MSB first, only
// the bits that the set of still-possible symbols leaves open; the table
// ends by repeating its last symbol) and the range coder follow libbsc's
// qlfc.cpp (absent from the reference tree) and are pinned by the libbsc
// archive under tests/golden. The rank/run model below is this library's
// own: libbsc's qlfc_model constants are not restated, so its coded bytes
// are not libbsc's.
//
// Model: every binary decision reads one adaptive probability from a context
// made of the current run's symbol and a short history state; its adaptation
// rate falls as 1/(n + 1.5) with the slot's update count n, down to 1/128.
#include "qlfc_coder.h"
#include "rangecoder.h"

#include <algorithm>
#include <cstring>
#include <memory>
#include <type_traits>

namespace qz {
namespace qlfc {

namespace {

// Adaptation rate after n updates, in 1/4096 units.
struct Rates {
    uint16_t rate[256];
    Rates()
    {
        for (int n = 0; n < 256; ++n) rate[n] = (uint16_t)(8192 / (2 * n + 3));
    }
};
// Built once at load time: the coding loops read it without a first-use
// guard on every binary decision.
const Rates g_rates;

struct Pred {
    uint16_t p;  // P(bit == 0), 16-bit
    uint16_t n;  // updates so far, saturating at kLimit
};

constexpr int kLimit = 127;

inline void upd(Pred& q, int bit)
{
    const int target = bit ? 0 : 65535;
    q.p = (uint16_t)(q.p + (((target - (int)q.p) * (int)g_rates.rate[q.n]) >> 12));
    if (q.n < kLimit) ++q.n;
}

inline int bsr(uint32_t v) { return 31 - __builtin_clz(v); }

// One binary decision coded with the probability held in q (12 bits for the
// range coder, kept inside [1, 4095]).
template <class IO>
inline int cbit(IO& io, int bit, Pred& q)
{
    int p = q.p >> 4;
    p = p < 1 ? 1 : p;
    bit = io.code(bit, p);
    upd(q, bit);
    return bit;
}

// Context slots. Rank v (0 .. nsym-1) is coded as a zero test, the unary
// exponent e = floor(log2 v) and the e bits below the leading one; run length
// L >= 1 as a test L > 1, the unary exponent of L and its bits below the
// leading one (the top three by tree node, the rest by position).
struct Model {
    Pred rankZero[256][64];      // symbol x (3 recent zero-test flags, rank class)
    Pred rankExp[256][8][8];     // symbol x rank class x exponent bit
    Pred rankMant[256][8][8];    // symbol x exponent x tree node 1..7
    Pred rankMantHi[8][128];     // exponent x tree node 8..127 (large alphabets)
    Pred runGt1[256][128];       // symbol x (rank zero, 3 recent run flags, run class)
    Pred runExp[256][8][32];     // symbol x run class x exponent bit
    Pred runMant[256][32][8];    // symbol x exponent x tree node 1..7
    Pred runLow[32][32];         // exponent x bit position
};
static_assert(std::is_trivially_copyable<Model>::value && sizeof(Model) % sizeof(Pred) == 0,
              "the model is a flat array of slots");

std::unique_ptr<Model> new_model()
{
    std::unique_ptr<Model> m(new Model);
    Pred* p = reinterpret_cast<Pred*>(m.get());
    std::fill(p, p + sizeof(Model) / sizeof(Pred), Pred{32768, 0});
    return m;
}

struct RunState {
    int nsym = 0, maxE = 0, maxLE = 1;
    uint32_t flagsV = 0, flagsL = 0;  // recent zero-test / length flags, newest in bit 0 (shifts wrap)
    uint8_t rankHist[256];  // per symbol: class of its last rank (0: zero, else min(e + 1, 7))
    uint8_t runHist[256];   // per symbol: class of its last run length
    RunState(int nsym_, int n)
    {
        nsym = nsym_;
        maxE = nsym > 1 ? bsr((uint32_t)(nsym - 1)) : 0;
        maxLE = n > 1 ? bsr((uint32_t)n) : 1;
        memset(rankHist, 0, sizeof(rankHist));
        memset(runHist, 0, sizeof(runHist));
    }
};

// Codes one run of symbol c: its rank (where c moves after the run) and its
// length. The encoder passes rank/len in; the decoder gets them out.
// Returns false on an impossible decoded value.
template <class IO>
inline bool code_run(IO& io, Model& M, RunState& st, int c, int& rank, uint32_t& len)
{
    int v = 0;
    if (st.nsym > 1) {
        if (IO::kEnc) v = rank ? rank - 1 : st.nsym - 1;
        const int rh = st.rankHist[c];
        const int nz = cbit(io, v != 0, M.rankZero[c][((st.flagsV & 7) << 3) | rh]);
        int e = 0;
        if (nz) {
            const int ev = IO::kEnc ? bsr((uint32_t)v) : 0;
            int k = 0;
            for (; k < st.maxE; ++k)
                if (!cbit(io, IO::kEnc ? (k < ev) : 0, M.rankExp[c][rh][k])) break;
            e = k;
            int node = 1;
            for (int b = e - 1; b >= 0; --b) {
                const int bit = IO::kEnc ? ((v >> b) & 1) : 0;
                Pred& q = node < 8 ? M.rankMant[c][e][node] : M.rankMantHi[e][node];
                node = node * 2 + cbit(io, bit, q);
            }
            v = node;
        }
        if (v > st.nsym - 1) return false;
        st.rankHist[c] = (uint8_t)(nz ? std::min(e + 1, 7) : 0);
        st.flagsV = (st.flagsV << 1) | (uint32_t)nz;
        if (!IO::kEnc) rank = (v == st.nsym - 1) ? 0 : v + 1;
    } else if (!IO::kEnc) {
        rank = 0;
    }

    const int rh = st.runHist[c];
    const int ctxL = ((v == 0) << 6) | ((st.flagsL & 7) << 3) | rh;
    const int gt1 = cbit(io, IO::kEnc ? (len > 1) : 0, M.runGt1[c][ctxL]);
    uint32_t L = 1;
    int e = 0;
    if (gt1) {
        const int ev = IO::kEnc ? bsr(len) : 0;
        int k = 1;
        for (; k < st.maxLE; ++k)
            if (!cbit(io, IO::kEnc ? (k < ev) : 0, M.runExp[c][rh][k])) break;
        e = k;
        int node = 1;
        for (int b = e - 1; b >= 0; --b) {
            const int bit = IO::kEnc ? (int)((len >> b) & 1u) : 0;
            int got;
            if (e - 1 - b < 3) {
                got = cbit(io, bit, M.runMant[c][e][node]);
                node = node * 2 + got;
            } else {
                got = cbit(io, bit, M.runLow[e][b]);
            }
            L = L * 2 + (uint32_t)got;
        }
    }
    st.runHist[c] = (uint8_t)(gt1 ? std::min(e + 1, 7) : 0);
    st.flagsL = (st.flagsL << 1) | (uint32_t)gt1;
    if (!IO::kEnc) len = L;
    return true;
}

// Candidate-set table coding (libbsc-compatible). The encoder walks
// table[0..nsym) then the repeated last symbol; the decoder rebuilds it.
template <class IO>
inline int code_table(IO& io, uint8_t* table, int nsym_in)
{
    uint8_t used[256];
    memset(used, 0, sizeof(used));
    int prev = -1, count = 0;
    for (int rank = 0; rank < 256; ++rank) {
        int cur = 0;
        if (IO::kEnc) cur = rank < nsym_in ? table[rank] : table[nsym_in - 1];
        int built = 0;
        for (int bit = 7; bit >= 0; --bit) {
            bool b0 = false, b1 = false;
            const int hi = (IO::kEnc ? cur : built) >> (bit + 1);
            for (int c = 0; c < 256; ++c) {
                if ((c == prev || !used[c]) && (c >> (bit + 1)) == hi) {
                    if (c & (1 << bit)) b1 = true; else b0 = true;
                    if (b0 && b1) break;
                }
            }
            int b;
            if (b0 && b1) b = io.code(IO::kEnc ? ((cur >> bit) & 1) : 0, 2048);
            else b = b1 ? 1 : 0;
            built |= b << bit;
        }
        if (built == prev) break;
        table[rank] = (uint8_t)built;
        prev = built;
        used[built] = 1;
        ++count;
    }
    return count;
}

}  // namespace

int num_segments(int n)
{
    // Fitted to the libbsc archive under tests/golden (2,904,234 B -> 6,
    // 4,450,000 and 5,050,000 B -> 8 segments); one segment up to 512 KiB.
    int nb = (int)(((int64_t)n + 524287) >> 19);
    return nb < 1 ? 1 : (nb > kMaxSegments ? kMaxSegments : nb);
}

void split_segments(const uint8_t* in, int n, int nb, int* start, int* size)
{
    int changes = 0;
    for (int i = 1; i < n; i += 32)
        if (in[i] != in[i - 1]) ++changes;
    if (changes > nb) {
        const int per = changes / nb;
        start[0] = 0;
        changes = 0;
        for (int id = 0, i = 1; i < n; i += 32) {
            if (in[i] != in[i - 1]) {
                if (++changes == per) {
                    changes = 0;
                    size[id] = i - start[id];
                    ++id;
                    start[id] = i;
                    if (id == nb - 1) break;
                }
            }
        }
        size[nb - 1] = n - start[nb - 1];
    } else {
        for (int p = 0; p < nb; ++p) {
            start[p] = (n / nb) * p;
            size[p] = (p != nb - 1) ? n / nb : n - (n / nb) * (nb - 1);
        }
    }
}

uint32_t front_end(const uint8_t* in, int n, uint8_t* sym, uint8_t* rank, uint32_t* len, uint8_t* table,
                   int* nsym)
{
    uint32_t runs = 0;
    for (int i = 0; i < n;) {
        int j = i + 1;
        while (j < n && in[j] == in[i]) ++j;
        sym[runs] = in[i];
        len[runs] = (uint32_t)(j - i);
        ++runs;
        i = j;
    }
    uint8_t list[256];
    int L = 0;
    for (int64_t t = (int64_t)runs - 1; t >= 0; --t) {
        const uint8_t c = sym[t];
        int p = 0;
        while (p < L && list[p] != c) ++p;
        rank[t] = (uint8_t)p;
        if (p == L) ++L;
        memmove(list + 1, list, (size_t)p);
        list[0] = c;
    }
    memcpy(table, list, (size_t)L);
    *nsym = L;
    return runs;
}

int encode_segment(const SegRuns& R, int n, uint8_t* out, int cap)
{
    if (n < 1 || R.nsym < 1) return -1;
    RangeEncoder enc;
    enc.init(out, cap);
    enc.put_word((uint32_t)n);
    uint8_t table[256];
    memcpy(table, R.table, (size_t)R.nsym);
    code_table(enc, table, R.nsym);
    std::unique_ptr<Model> M = new_model();
    RunState st(R.nsym, n);
    for (uint32_t i = 0; i < R.count; ++i) {
        if (enc.eob()) return -3;
        int rank = R.rank[i];
        uint32_t len = R.len[i];
        code_run(enc, *M, st, R.sym[i], rank, len);
    }
    const int r = enc.finish();
    return r < 0 ? -3 : r;
}

int read_segment_header(const uint8_t* in, int in_size, int* n, uint8_t* table, int* nsym)
{
    RangeDecoder dec;
    dec.init(in, in_size);
    *n = (int)dec.get_word();
    *nsym = code_table(dec, table, 0);
    return dec.overrun() ? -5 : 0;
}

int decode_segment(const uint8_t* in, int in_size, uint8_t* out, int n)
{
    RangeDecoder dec;
    dec.init(in, in_size);
    const uint32_t word = dec.get_word();
    if ((int64_t)word != (int64_t)n) return -6;
    uint8_t list[256] = {};
    const int nsym = code_table(dec, list, 0);
    if (nsym < 1) return -6;
    std::unique_ptr<Model> M = new_model();
    RunState st(nsym, n);
    int pos = 0;
    // Up to 8 symbols, the list lives in one word (byte i = list[i]) and a
    // run's move to position `rank` is done with masks; short runs are
    // written as one 16-byte store when 16 bytes remain. Both avoid branches
    // whose trip counts depend on decoded values.
    const bool small = nsym <= 8;
    uint64_t W = 0;
    if (small) memcpy(&W, list, 8);
    while (pos < n) {
        const int c = small ? (int)(W & 0xff) : list[0];
        int rank = 0;
        uint32_t len = 0;
        if (!code_run(dec, *M, st, c, rank, len)) return -6;
        if (len > (uint32_t)(n - pos) || dec.overrun()) return -6;
        if (n - pos >= 16) {
            uint8_t fill[16];
            memset(fill, c, 16);
            memcpy(out + pos, fill, 16);
            if (len > 16) memset(out + pos + 16, c, len - 16);
        } else {
            memset(out + pos, c, len);
        }
        pos += (int)len;
        if (rank > 0) {
            if (rank >= nsym) return -6;
            if (small) {
                const int sh = 8 * rank;
                const uint64_t low = (W >> 8) & ((1ull << sh) - 1);
                const uint64_t high = sh + 8 >= 64 ? 0 : W & ~((1ull << (sh + 8)) - 1);
                W = low | ((uint64_t)c << sh) | high;
            } else {
                for (int i = 0; i < rank; ++i) list[i] = list[i + 1];
                list[rank] = (uint8_t)c;
            }
        } else if (pos != n) {
            return -6;
        }
    }
    return 0;
}

}  // namespace qlfc
}  // namespace qz
