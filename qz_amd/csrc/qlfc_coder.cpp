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
// Model: every binary decision mixes three adaptive probabilities (a static
// one per decision node, one selected by recent rank/run history, one by the
// current symbol) in the logistic domain with per-context weights.
#include "qlfc_coder.h"
#include "rangecoder.h"

#include <algorithm>
#include <cstring>
#include <memory>

namespace qz {
namespace qlfc {

namespace {

// Logistic 4096/(1+exp(-x/256)) sampled at x = -2048..2048 step 128.
const int kSquashPts[33] = {1,    2,    4,    6,    10,   17,   27,   45,   74,   120,  194,
                            311,  488,  747,  1102, 1546, 2048, 2550, 2994, 3349, 3608, 3785,
                            3902, 3976, 4022, 4051, 4069, 4079, 4086, 4090, 4092, 4094, 4095};

struct Tables {
    int16_t squash[4096];  // index x + 2048
    int16_t stretch[4096];
    uint16_t rate[256];    // adaptation rate after n updates, 1/4096 units
    Tables()
    {
        for (int d = -2048; d < 2048; ++d) {
            const int x = d < -2047 ? -2047 : d;
            const int w = x & 127, i = (x >> 7) + 16;
            squash[d + 2048] = (int16_t)((kSquashPts[i] * (128 - w) + kSquashPts[i + 1] * w + 64) >> 7);
        }
        int pi = 0;
        for (int x = -2047; x <= 2047; ++x) {
            const int v = squash[x + 2048];
            for (int p = pi; p <= v; ++p) stretch[p] = (int16_t)x;
            pi = v + 1;
        }
        for (int p = pi; p < 4096; ++p) stretch[p] = 2047;
        for (int n = 0; n < 256; ++n) rate[n] = (uint16_t)(8192 / (2 * n + 3));
    }
};
// Built once at load time: the coding loops read it without a
// first-use guard on every binary decision.
const Tables g_tables;
inline const Tables& tab() { return g_tables; }

struct Pred {
    uint16_t p;  // P(bit == 0), 16-bit
    uint16_t n;
};

struct Mix {
    int32_t w[3];
};

constexpr int kLimStatic = 60, kLimCtx = 30, kLr = 6;

inline void upd(Pred& q, int bit, int lim)
{
    const int target = bit ? 0 : 65535;
    q.p = (uint16_t)(q.p + (((target - (int)q.p) * (int)tab().rate[q.n]) >> 12));
    if (q.n < lim) ++q.n;
}

inline int bsr(uint32_t v) { return 31 - __builtin_clz(v); }

template <class IO>
inline int cbit(IO& io, int bit, Pred& a, Pred& b, Pred& c, Mix& m)
{
    const Tables& t = tab();
    const int s0 = t.stretch[a.p >> 4], s1 = t.stretch[b.p >> 4], s2 = t.stretch[c.p >> 4];
    int d = (int)(((int64_t)m.w[0] * s0 + (int64_t)m.w[1] * s1 + (int64_t)m.w[2] * s2) >> 16);
    d = d < -2047 ? -2047 : (d > 2047 ? 2047 : d);
    const int pm = t.squash[d + 2048];
    bit = io.code(bit, pm);
    upd(a, bit, kLimStatic);
    upd(b, bit, kLimCtx);
    upd(c, bit, kLimCtx);
    const int err = ((bit ? 0 : 4095) - pm) * kLr;
    m.w[0] += (s0 * err) >> 14;
    m.w[1] += (s1 * err) >> 14;
    m.w[2] += (s2 * err) >> 14;
    return bit;
}

template <class IO>
inline int dbit(IO& io, int bit, Pred& q)
{
    int p = q.p >> 4;
    p = p < 1 ? 1 : p;
    bit = io.code(bit, p);
    upd(q, bit, kLimStatic);
    return bit;
}

struct Model {
    // rank: zero test, unary exponent, mantissa tree
    Pred r0s, r0t[64], r0c[256];
    Mix r0m[256];
    Pred res[8], ret[64][8], rec[256][8];
    Mix rem[8][8];
    Pred rms[8][128], rmt[8][8][128], rmc[256][8][8];
    Mix rmm[8];
    // run length: >1 test, unary exponent, top mantissa bits, low bits
    Pred l0s, l0t[128], l0c[256];
    Mix l0m[256];
    Pred les[32], lett[128][32], lec[256][32];
    Mix lem[32];
    Pred lms[32][8], lmt[8][32][8], lmc[256][32][8];
    Mix lmm[32];
    Pred lbs[32][32];
};

const Model& model_template()
{
    static const Model* m = [] {
        Model* x = new Model;
        Pred* p = reinterpret_cast<Pred*>(x);
        // every member is an array of Pred or Mix; initialise by type
        auto initp = [](Pred* a, size_t n) {
            for (size_t i = 0; i < n; ++i) a[i] = Pred{32768, 0};
        };
        auto initm = [](Mix* a, size_t n) {
            for (size_t i = 0; i < n; ++i) a[i] = Mix{{21845, 21845, 21845}};
        };
        (void)p;
        initp(&x->r0s, 1);
        initp(x->r0t, 64);
        initp(x->r0c, 256);
        initm(x->r0m, 256);
        initp(x->res, 8);
        initp(&x->ret[0][0], 64 * 8);
        initp(&x->rec[0][0], 256 * 8);
        initm(&x->rem[0][0], 64);
        initp(&x->rms[0][0], 8 * 128);
        initp(&x->rmt[0][0][0], 8 * 8 * 128);
        initp(&x->rmc[0][0][0], 256 * 8 * 8);
        initm(x->rmm, 8);
        initp(&x->l0s, 1);
        initp(x->l0t, 128);
        initp(x->l0c, 256);
        initm(x->l0m, 256);
        initp(x->les, 32);
        initp(&x->lett[0][0], 128 * 32);
        initp(&x->lec[0][0], 256 * 32);
        initm(x->lem, 32);
        initp(&x->lms[0][0], 32 * 8);
        initp(&x->lmt[0][0][0], 8 * 32 * 8);
        initp(&x->lmc[0][0][0], 256 * 32 * 8);
        initm(x->lmm, 32);
        initp(&x->lbs[0][0], 32 * 32);
        return x;
    }();
    return *m;
}

struct RunState {
    int nsym = 0, maxE = 0, maxLE = 1;
    int flagsV = 0, flagsL = 0;
    uint8_t rankHist[256];
    uint8_t runHist[256];
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
        const int ctxR = ((st.flagsV & 7) << 3) | rh;
        const int nz = cbit(io, v != 0, M.r0s, M.r0t[ctxR], M.r0c[c], M.r0m[c]);
        int e = 0;
        if (nz) {
            const int ev = IO::kEnc ? bsr((uint32_t)v) : 0;
            int k = 0;
            for (; k < st.maxE; ++k)
                if (!cbit(io, IO::kEnc ? (k < ev) : 0, M.res[k], M.ret[ctxR][k], M.rec[c][k], M.rem[rh][k])) break;
            e = k;
            int node = 1;
            const int h3 = st.flagsV & 7;
            for (int b = e - 1; b >= 0; --b) {
                const int bit = IO::kEnc ? ((v >> b) & 1) : 0;
                const int got = cbit(io, bit, M.rms[e][node], M.rmt[h3][e][node], M.rmc[c][e][node < 8 ? node : 0],
                                     M.rmm[e]);
                node = node * 2 + got;
            }
            v = node;
        }
        if (v > st.nsym - 1) return false;
        st.rankHist[c] = (uint8_t)(nz ? std::min(e + 1, 7) : 0);
        st.flagsV = (st.flagsV << 1) | nz;
        if (!IO::kEnc) rank = (v == st.nsym - 1) ? 0 : v + 1;
    } else if (!IO::kEnc) {
        rank = 0;
    }

    const int rh = st.runHist[c];
    const int ctxL = ((v == 0) << 6) | ((st.flagsL & 7) << 3) | rh;
    const int gt1 = cbit(io, IO::kEnc ? (len > 1) : 0, M.l0s, M.l0t[ctxL], M.l0c[c], M.l0m[c]);
    uint32_t L = 1;
    int e = 0;
    if (gt1) {
        const int ev = IO::kEnc ? bsr(len) : 0;
        int k = 1;
        for (; k < st.maxLE; ++k)
            if (!cbit(io, IO::kEnc ? (k < ev) : 0, M.les[k], M.lett[ctxL][k], M.lec[c][k], M.lem[k])) break;
        e = k;
        const int h3 = st.flagsL & 7;
        int node = 1;
        for (int b = e - 1; b >= 0; --b) {
            const int bit = IO::kEnc ? (int)((len >> b) & 1u) : 0;
            int got;
            if (e - 1 - b < 3) {
                got = cbit(io, bit, M.lms[e][node], M.lmt[h3][e][node], M.lmc[c][e][node], M.lmm[e]);
                node = node * 2 + got;
            } else {
                got = dbit(io, bit, M.lbs[e][b]);
            }
            L = L * 2 + (uint32_t)got;
        }
    }
    st.runHist[c] = (uint8_t)(gt1 ? std::min(e + 1, 7) : 0);
    st.flagsL = (st.flagsL << 1) | gt1;
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
    std::unique_ptr<Model> M(new Model(model_template()));
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
    uint8_t list[256];
    const int nsym = code_table(dec, list, 0);
    if (nsym < 1) return -6;
    std::unique_ptr<Model> M(new Model(model_template()));
    RunState st(nsym, n);
    int pos = 0;
    while (pos < n) {
        const int c = list[0];
        int rank = 0;
        uint32_t len = 0;
        if (!code_run(dec, *M, st, c, rank, len)) return -6;
        if (len > (uint32_t)(n - pos) || dec.overrun()) return -6;
        memset(out + pos, c, len);
        pos += (int)len;
        if (rank > 0) {
            if (rank >= nsym) return -6;
            memmove(list, list + 1, (size_t)rank);
            list[rank] = (uint8_t)c;
        } else if (pos != n) {
            return -6;
        }
    }
    return 0;
}

}  // namespace qlfc
}  // namespace qz
