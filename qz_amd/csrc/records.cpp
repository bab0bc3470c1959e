// FASTQ records <-> the three streams of a .qz archive (host): the record
// layout of qz's records_to_streams (crates/qz-lib/src/compression/
// mod.rs:280-320) for encoding 0 with lossless qualities, and the offset
// scans and record assembly of its decoder (decompress_impl.rs:444-636).
// Records are read as qz's FastqReader reads them (io/fastq.rs:141-197):
// four lines, the third starting with '+', trailing '\r's dropped, the
// quality as long as the sequence; the id line is taken as it is.
//
// Streams, per record in order:
//   header   varint(len) + id (the id keeps its '@')
//   sequence the bases, preceded by varint(len) unless every read has the
//            same length (const_seq)
//   quality  min(q - 33, 127) (0 below '!') as 7-bit codes packed least
//            significant bit first, flushed to a byte per record
//            (columnar.rs:103-134), preceded by varint(len) unless every
//            record has the same quality length (const_qual)
// varint: 7 bits per byte, least significant group first (dna_utils.rs:
// 285-311).
#include "common.h"
#include "../../include/qzbsc.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr int kQualBits = 7;

inline long long packed_len(long long nq) { return (nq * kQualBits + 7) / 8; }

inline int varint_len(unsigned long long v)
{
    int k = 1;
    while (v >= 0x80) {
        v >>= 7;
        ++k;
    }
    return k;
}

inline unsigned char* put_varint(unsigned char* p, unsigned long long v)
{
    while (v >= 0x80) {
        *p++ = (unsigned char)(v | 0x80);
        v >>= 7;
    }
    *p++ = (unsigned char)v;
    return p;
}

// Reads a varint at data[*pos]; false when it runs past `size` or is too long.
inline bool get_varint(const unsigned char* data, long long size, long long* pos, unsigned long long* v)
{
    unsigned long long x = 0;
    for (int shift = 0; shift <= 63; shift += 7) {
        if (*pos >= size) return false;
        const unsigned char b = data[(*pos)++];
        x |= (unsigned long long)(b & 0x7f) << shift;
        if (!(b & 0x80)) {
            *v = x;
            return true;
        }
    }
    return false;
}

inline unsigned char* pack_quals(unsigned char* p, const unsigned char* q, long long n)
{
    unsigned int buf = 0;
    int nb = 0;
    for (long long i = 0; i < n; ++i) {
        int c = (int)q[i] - 33;
        c = c < 0 ? 0 : (c > 127 ? 127 : c);
        buf |= (unsigned int)c << nb;
        nb += kQualBits;
        while (nb >= 8) {
            *p++ = (unsigned char)buf;
            buf >>= 8;
            nb -= 8;
        }
    }
    if (nb > 0) *p++ = (unsigned char)buf;
    return p;
}

inline unsigned char* unpack_quals(unsigned char* o, const unsigned char* p, long long n)
{
    unsigned int buf = 0;
    int nb = 0;
    for (long long i = 0; i < n; ++i) {
        if (nb < kQualBits) {
            buf |= (unsigned int)*p++ << nb;
            nb += 8;
        }
        *o++ = (unsigned char)((buf & 0x7f) + 33);
        buf >>= kQualBits;
        nb -= kQualBits;
    }
    return o;
}

struct Line {
    long long s, e;  // [s, e) without the newline
};

// The next line of the text at *pos; false at the end.
inline bool next_line(const unsigned char* t, long long size, long long* pos, Line* l)
{
    if (*pos >= size) return false;
    const void* nl = memchr(t + *pos, '\n', (size_t)(size - *pos));
    l->s = *pos;
    l->e = nl ? (long long)((const unsigned char*)nl - t) : size;
    *pos = nl ? l->e + 1 : size;
    return true;
}

// Line ends as qz's reader takes them (io/fastq.rs:132-139): trailing '\r's
// are dropped, so CRLF text reads as LF text.
inline void trim_cr(const unsigned char* t, Line* l)
{
    while (l->e > l->s && t[l->e - 1] == '\r') --l->e;
}

// Texts and stream sets below this size convert on one thread.
constexpr long long kParallelBytes = 4ll << 20;

int conversion_threads(long long bytes)
{
    if (bytes < kParallelBytes) return 1;
    const long long hw = (long long)std::max(1u, std::thread::hardware_concurrency());
    return (int)std::max(1ll, std::min(hw, bytes / (kParallelBytes / 2)));
}

// Runs fn(t) for t in [0, T) on T threads (none when T is 0).
template <class Fn>
void run_threads(int T, Fn fn)
{
    if (T <= 0) return;
    if (T == 1) {
        fn(0);
        return;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(fn, t);
    fn(0);
    for (auto& x : th) x.join();
}

// Record totals of a stretch of FASTQ text.
struct Scan {
    long long rec = 0;
    long long h = 0;               // header stream bytes
    long long s_var = 0, s_raw = 0;  // sequence bytes with / without the length varints
    long long q_var = 0, q_raw = 0;  // packed quality bytes with / without them
    long long s1 = -1, q1 = -1;    // first record's lengths
    bool s_same = true, q_same = true;
};

// Parses the records of fastq[start, stop) (the last one must end exactly at
// stop), adding them to sc. With out pointers, also writes them.
bool parse_records(const unsigned char* fastq, long long size, long long start, long long stop, Scan* sc,
                   unsigned char* hp, unsigned char* sp, unsigned char* qp, bool cs, bool cq)
{
    long long pos = start;
    Line h, s, p, q;
    while (pos < stop) {
        if (!next_line(fastq, size, &pos, &h) || !next_line(fastq, size, &pos, &s) ||
            !next_line(fastq, size, &pos, &p) || !next_line(fastq, size, &pos, &q))
            return false;  // a record of fewer than four lines
        if (p.e == p.s || fastq[p.s] != '+') return false;
        trim_cr(fastq, &h);
        trim_cr(fastq, &s);
        trim_cr(fastq, &q);
        const long long hl = h.e - h.s, sl = s.e - s.s, ql = q.e - q.s;
        if (ql != sl) return false;  // io/fastq.rs:183-190
        sc->h += varint_len((unsigned long long)hl) + hl;
        sc->s_raw += sl;
        sc->s_var += varint_len((unsigned long long)sl) + sl;
        sc->q_raw += packed_len(ql);
        sc->q_var += varint_len((unsigned long long)ql) + packed_len(ql);
        if (sc->rec == 0) {
            sc->s1 = sl;
            sc->q1 = ql;
        }
        sc->s_same = sc->s_same && sl == sc->s1;
        sc->q_same = sc->q_same && ql == sc->q1;
        ++sc->rec;
        if (hp) {
            hp = put_varint(hp, (unsigned long long)hl);
            memcpy(hp, fastq + h.s, (size_t)hl);
            hp += hl;
            if (!cs) sp = put_varint(sp, (unsigned long long)sl);
            memcpy(sp, fastq + s.s, (size_t)sl);
            sp += sl;
            if (!cq) qp = put_varint(qp, (unsigned long long)ql);
            qp = pack_quals(qp, fastq + q.s, ql);
        }
    }
    return pos == stop;
}

// Start of the first record at or after x: a line beginning with '@' whose
// second line below begins with '+'. A quality line beginning with '@' fails
// this test (two lines below it is a sequence line), so in valid FASTQ the
// match is a record start. Returns -1 if none is found within a few lines.
long long record_start(const unsigned char* t, long long size, long long x)
{
    if (x <= 0) return 0;
    long long pos = x;
    if (t[pos - 1] != '\n') {
        const void* nl = memchr(t + pos, '\n', (size_t)(size - pos));
        if (!nl) return -1;
        pos = (long long)((const unsigned char*)nl - t) + 1;
    }
    for (int tries = 0; tries < 8 && pos < size; ++tries) {
        long long p = pos;
        Line a, b, c;
        if (!next_line(t, size, &p, &a) || !next_line(t, size, &p, &b) || !next_line(t, size, &p, &c)) return -1;
        if (a.e > a.s && t[a.s] == '@' && c.e > c.s && t[c.s] == '+') return pos;
        long long q = pos;
        next_line(t, size, &q, &a);
        pos = q;
    }
    return -1;
}

// Splits the text into up to T stretches at record starts and parses them in
// parallel. Falls back to one stretch when the cuts cannot be confirmed: the
// result is always that of one sequential parse.
bool scan_text(const unsigned char* fastq, long long size, std::vector<long long>& cut, std::vector<Scan>& sc)
{
    const int T = conversion_threads(size);
    cut.assign(1, 0);
    for (int t = 1; t < T; ++t) {
        const long long c = record_start(fastq, size, size * t / T);
        if (c < 0) break;
        if (c > cut.back() && c < size) cut.push_back(c);
    }
    cut.push_back(size);
    const int K = (int)cut.size() - 1;
    sc.assign(K, Scan{});
    std::vector<char> ok(K, 0);
    run_threads(K, [&](int k) {
        ok[k] = parse_records(fastq, size, cut[k], cut[k + 1], &sc[k], nullptr, nullptr, nullptr, false, false);
    });
    if (std::all_of(ok.begin(), ok.end(), [](char v) { return v != 0; })) return true;
    if (K == 1) return false;
    cut.assign({0, size});
    sc.assign(1, Scan{});
    return parse_records(fastq, size, 0, size, &sc[0], nullptr, nullptr, nullptr, false, false);
}

}  // namespace

extern "C" {

QZBSC_API long long qzbsc_fastq_to_streams(const unsigned char* fastq, long long size, unsigned char* hdr,
                                           unsigned char* seq, unsigned char* qual, long long* info)
{
    if ((!fastq && size) || size < 0 || !info) return QZ_BAD_PARAMETER;
    std::vector<long long> cut;
    std::vector<Scan> sc;
    if (!scan_text(fastq, size, cut, sc)) return QZ_BAD_PARAMETER;
    // totals and the const lengths (mod.rs:329-354): no records or any
    // variation gives 0; an all-empty quality set is not recorded as constant
    Scan all;
    for (const Scan& x : sc) {
        if (!x.rec) continue;
        if (!all.rec) {
            all.s1 = x.s1;
            all.q1 = x.q1;
        }
        all.s_same = all.s_same && x.s_same && x.s1 == all.s1;
        all.q_same = all.q_same && x.q_same && x.q1 == all.q1;
        all.rec += x.rec;
        all.h += x.h;
        all.s_var += x.s_var;
        all.s_raw += x.s_raw;
        all.q_var += x.q_var;
        all.q_raw += x.q_raw;
    }
    const long long cs = (all.rec && all.s_same) ? all.s1 : 0;
    const long long cq = (all.rec && all.q_same) ? all.q1 : 0;
    const bool write = hdr && seq && qual;
    if (!write) {
        info[0] = all.rec;
        info[1] = cs;
        info[2] = cq;
        info[3] = all.h;
        info[4] = cs ? all.s_raw : all.s_var;
        info[5] = cq ? all.q_raw : all.q_var;
        return 0;
    }
    // the write pass takes the record count and const lengths of the sizing
    // pass; anything else does not match the buffers' sizes
    if (info[0] != all.rec || info[1] != cs || info[2] != cq) return QZ_BAD_PARAMETER;
    const int K = (int)sc.size();
    std::vector<long long> ho(K + 1, 0), so(K + 1, 0), qo(K + 1, 0);
    for (int k = 0; k < K; ++k) {
        ho[k + 1] = ho[k] + sc[k].h;
        so[k + 1] = so[k] + (cs ? sc[k].s_raw : sc[k].s_var);
        qo[k + 1] = qo[k] + (cq ? sc[k].q_raw : sc[k].q_var);
    }
    std::vector<char> ok(K, 0);
    run_threads(K, [&](int k) {
        Scan tmp;
        ok[k] = parse_records(fastq, size, cut[k], cut[k + 1], &tmp, hdr + ho[k], seq + so[k], qual + qo[k], cs != 0,
                              cq != 0);
    });
    return std::all_of(ok.begin(), ok.end(), [](char v) { return v != 0; }) ? 0 : QZ_BAD_PARAMETER;
}

QZBSC_API long long qzbsc_streams_to_fastq(const unsigned char* hs, long long hlen, const unsigned char* ss,
                                           long long slen, const unsigned char* qs, long long qlen, long long n,
                                           long long const_seq, long long const_qual, unsigned char* out,
                                           long long cap)
{
    if (n < 0 || hlen < 0 || slen < 0 || qlen < 0 || const_seq < 0 || const_qual < 0) return QZ_BAD_PARAMETER;
    const bool has_q = qlen > 0;
    // Validating scan; every K-th record's stream and output positions are
    // kept so that the records can then be written in parallel.
    struct Pos {
        long long h, s, q, o;
    };
    const int T = conversion_threads(hlen + slen + qlen);
    const long long K = std::max(1ll, (n + T - 1) / T);
    std::vector<Pos> cp;
    long long hp = 0, sp = 0, qp = 0, o = 0;
    for (long long i = 0; i < n; ++i) {
        if (i % K == 0) cp.push_back({hp, sp, qp, o});
        unsigned long long hl, sl, ql = 0;
        if (!get_varint(hs, hlen, &hp, &hl) || hl > (unsigned long long)(hlen - hp)) return QZ_UNEXPECTED_EOB;
        if (const_seq) {
            sl = (unsigned long long)const_seq;
        } else if (!get_varint(ss, slen, &sp, &sl)) {
            return QZ_UNEXPECTED_EOB;
        }
        if (sl > (unsigned long long)(slen - sp)) return QZ_UNEXPECTED_EOB;
        long long pl = 0;
        if (has_q) {
            if (const_qual) {
                ql = (unsigned long long)const_qual;
            } else if (!get_varint(qs, qlen, &qp, &ql)) {
                return QZ_UNEXPECTED_EOB;
            }
            if (ql > (1ull << 40)) return QZ_DATA_CORRUPT;
            pl = packed_len((long long)ql);
            if (pl > qlen - qp) return QZ_UNEXPECTED_EOB;
        }
        o += (long long)(hl + sl + ql) + 5;  // id \n seq \n + \n qual \n
        hp += (long long)hl;
        sp += (long long)sl;
        qp += pl;
    }
    if (!out) return o;
    if (o > cap) return QZ_BAD_PARAMETER;
    const int P = (int)cp.size();
    run_threads(P, [&](int t) {
        long long h = cp[t].h, s = cp[t].s, q = cp[t].q;
        unsigned char* w = out + cp[t].o;
        const long long end = std::min(n, (t + 1) * K);
        for (long long i = t * K; i < end; ++i) {
            unsigned long long hl, sl, ql = 0;
            get_varint(hs, hlen, &h, &hl);  // validated by the scan
            if (const_seq) sl = (unsigned long long)const_seq;
            else get_varint(ss, slen, &s, &sl);
            if (has_q) {
                if (const_qual) ql = (unsigned long long)const_qual;
                else get_varint(qs, qlen, &q, &ql);
            }
            memcpy(w, hs + h, hl);
            w += hl;
            *w++ = '\n';
            memcpy(w, ss + s, sl);
            w += sl;
            *w++ = '\n';
            *w++ = '+';
            *w++ = '\n';
            if (has_q) w = unpack_quals(w, qs + q, (long long)ql);
            *w++ = '\n';
            h += (long long)hl;
            s += (long long)sl;
            q += packed_len((long long)ql);
        }
    });
    return o;
}

}  // extern "C"
