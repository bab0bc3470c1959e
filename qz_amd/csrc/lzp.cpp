// LZP preprocessing (host). The classic context-prediction scheme that
// libbsc applies before the BWT when bsc_compress is given lzpHashSize and
// lzpMinLen: the four bytes before a position select a slot of a table of
// 2^hash_bits positions, the slot predicts that the text continues as it
// did after the position it holds, and a prediction that holds for at least
// min_len bytes is sent as one token.
//
// Stream: [escape byte E] then tokens. E is the rarest byte value of the
// block (the smallest such value on ties), so a block that lacks some byte
// never escapes a literal.
//   byte b != E          literal b
//   E, 0                 literal E
//   E, varint(v >= 1)    match of length v - 1 + min_len, copied from the
//                        predicted position
// varint: 7 bits per byte, least significant group first, high bit set on
// every byte but the last. Table slots are updated at every token start at
// position >= 4 (never inside a match), identically on both sides.
#include "lzp.h"

#include "common.h"

#include <cstring>
#include <vector>

namespace qz {
namespace lzp {

bool params_ok(int hash_bits, int min_len) { return hash_bits >= 10 && hash_bits <= 28 && min_len >= 4 && min_len <= 255; }

static inline uint32_t slot(const uint8_t* p4, int hash_bits)
{
    uint32_t ctx;
    memcpy(&ctx, p4, 4);
    return (ctx * 2654435761u) >> (32 - hash_bits);
}

static inline int put_varint(uint8_t* out, uint32_t v)
{
    int k = 0;
    while (v >= 0x80) {
        out[k++] = (uint8_t)(v | 0x80);
        v >>= 7;
    }
    out[k++] = (uint8_t)v;
    return k;
}

int encode(const uint8_t* in, int n, uint8_t* out, int cap, int hash_bits, int min_len)
{
    if (!params_ok(hash_bits, min_len) || n < 0 || !in || !out) return QZ_BAD_PARAMETER;
    if (n == 0 || cap < 1) return QZ_NOT_COMPRESSIBLE;
    const int limit = cap < n ? cap : n;  // the result must be shorter than the input
    uint32_t freq[256] = {0};
    for (int i = 0; i < n; ++i) ++freq[in[i]];
    int esc = 0;
    for (int c = 1; c < 256; ++c)
        if (freq[c] < freq[esc]) esc = c;
    std::vector<int32_t> table((size_t)1 << hash_bits, 0);
    int o = 0;
    out[o++] = (uint8_t)esc;
    int i = 0;
    while (i < n) {
        if (o + 7 > limit) return QZ_NOT_COMPRESSIBLE;
        if (i >= 4) {
            int32_t& t = table[slot(in + i - 4, hash_bits)];
            const int p = t;
            t = i;
            // the last byte of a minimum-length match first: most misses end there
            if (p > 0 && i + min_len <= n && in[p + min_len - 1] == in[i + min_len - 1]) {
                int len = 0;
                const int max = n - i;
                while (len < max && in[p + len] == in[i + len]) ++len;
                if (len >= min_len) {
                    out[o++] = (uint8_t)esc;
                    o += put_varint(out + o, (uint32_t)(len - min_len + 1));
                    i += len;
                    continue;
                }
            }
        }
        const uint8_t b = in[i++];
        out[o++] = b;
        if (b == esc) out[o++] = 0;
    }
    return o < limit ? o : QZ_NOT_COMPRESSIBLE;
}

int decode(const uint8_t* in, int m, uint8_t* out, int n, int hash_bits, int min_len)
{
    if (!params_ok(hash_bits, min_len) || n < 0 || m < 1 || !in || !out) return QZ_BAD_PARAMETER;
    const uint8_t esc = in[0];
    std::vector<int32_t> table((size_t)1 << hash_bits, 0);
    int r = 1, i = 0;
    while (i < n) {
        int p = 0;
        if (i >= 4) {
            int32_t& t = table[slot(out + i - 4, hash_bits)];
            p = t;
            t = i;
        }
        if (r >= m) return QZ_DATA_CORRUPT;
        const uint8_t b = in[r++];
        if (b != esc) {
            out[i++] = b;
            continue;
        }
        uint32_t v = 0;
        for (int shift = 0;; shift += 7) {
            if (r >= m || shift > 28) return QZ_DATA_CORRUPT;
            const uint8_t x = in[r++];
            v |= (uint32_t)(x & 0x7f) << shift;
            if (!(x & 0x80)) break;
        }
        if (v == 0) {
            out[i++] = esc;
            continue;
        }
        const uint64_t len = (uint64_t)v - 1 + (uint64_t)min_len;
        if (p == 0 || (uint64_t)i + len > (uint64_t)n) return QZ_DATA_CORRUPT;
        for (uint64_t k = 0; k < len; ++k) out[i + k] = out[p + k];  // p < i: overlapping copies repeat
        i += (int)len;
    }
    return r == m ? 0 : QZ_DATA_CORRUPT;
}

}  // namespace lzp
}  // namespace qz
