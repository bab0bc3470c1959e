// FASTQ records <-> the three streams of a .qz archive (host): the record
// layout of qz's records_to_streams (crates/qz-lib/src/compression/
// mod.rs:280-320) for encoding 0 with lossless qualities, and the offset
// scans and record assembly of its decoder (decompress_impl.rs:444-636).
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

#include <cstring>

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

}  // namespace

extern "C" {

QZBSC_API long long qzbsc_fastq_to_streams(const unsigned char* fastq, long long size, unsigned char* hdr,
                                           unsigned char* seq, unsigned char* qual, long long* info)
{
    if ((!fastq && size) || size < 0 || !info) return QZ_BAD_PARAMETER;
    const bool write = hdr && seq && qual;
    long long n = 0, hlen = 0, slen = 0, qlen = 0, cs = -1, cq = -1;
    if (write) {
        n = info[0];
        cs = info[1];
        cq = info[2];
    }
    unsigned char *hp = hdr, *sp = seq, *qp = qual;
    long long pos = 0, rec = 0;
    Line h, s, p, q;
    while (next_line(fastq, size, &pos, &h)) {
        if (!next_line(fastq, size, &pos, &s) || !next_line(fastq, size, &pos, &p) ||
            !next_line(fastq, size, &pos, &q))
            return QZ_BAD_PARAMETER;  // a record of fewer than four lines
        if (h.e == h.s || fastq[h.s] != '@' || p.e == p.s || fastq[p.s] != '+') return QZ_BAD_PARAMETER;
        const long long hl = h.e - h.s, sl = s.e - s.s, ql = q.e - q.s;
        if (!write) {
            hlen += varint_len((unsigned long long)hl) + hl;
            slen += varint_len((unsigned long long)sl) + sl;
            qlen += varint_len((unsigned long long)ql) + packed_len(ql);
            cs = (rec == 0) ? sl : (cs == sl ? cs : 0);
            cq = (rec == 0) ? ql : (cq == ql ? cq : 0);
        } else {
            if (rec >= n) return QZ_BAD_PARAMETER;
            hp = put_varint(hp, (unsigned long long)hl);
            memcpy(hp, fastq + h.s, (size_t)hl);
            hp += hl;
            if (!cs) sp = put_varint(sp, (unsigned long long)sl);
            memcpy(sp, fastq + s.s, (size_t)sl);
            sp += sl;
            if (!cq) qp = put_varint(qp, (unsigned long long)ql);
            qp = pack_quals(qp, fastq + q.s, ql);
        }
        ++rec;
    }
    if (write) {
        if (rec != n) return QZ_BAD_PARAMETER;
        return 0;
    }
    // const lengths (mod.rs:329-354): no records or any variation gives 0;
    // an all-empty quality set is not recorded as constant
    if (rec == 0) cs = cq = 0;
    if (cq < 0) cq = 0;
    if (cs > 0) slen -= rec * varint_len((unsigned long long)cs);
    if (cq > 0) qlen -= rec * varint_len((unsigned long long)cq);
    info[0] = rec;
    info[1] = cs;
    info[2] = cq;
    info[3] = hlen;
    info[4] = slen;
    info[5] = qlen;
    return 0;
}

QZBSC_API long long qzbsc_streams_to_fastq(const unsigned char* hs, long long hlen, const unsigned char* ss,
                                           long long slen, const unsigned char* qs, long long qlen, long long n,
                                           long long const_seq, long long const_qual, unsigned char* out,
                                           long long cap)
{
    if (n < 0 || hlen < 0 || slen < 0 || qlen < 0 || const_seq < 0 || const_qual < 0) return QZ_BAD_PARAMETER;
    const bool has_q = qlen > 0;
    long long hp = 0, sp = 0, qp = 0, o = 0;
    for (long long i = 0; i < n; ++i) {
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
        const long long rl = (long long)(hl + sl + ql) + 5;  // id \n seq \n + \n qual \n
        if (out) {
            if (o + rl > cap) return QZ_BAD_PARAMETER;
            unsigned char* w = out + o;
            memcpy(w, hs + hp, hl);
            w += hl;
            *w++ = '\n';
            memcpy(w, ss + sp, sl);
            w += sl;
            *w++ = '\n';
            *w++ = '+';
            *w++ = '\n';
            if (has_q) w = unpack_quals(w, qs + qp, (long long)ql);
            *w++ = '\n';
        }
        o += rl;
        hp += (long long)hl;
        sp += (long long)sl;
        qp += pl;
    }
    return o;
}

}  // extern "C"
