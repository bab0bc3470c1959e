/*
 * ORACLE — test infrastructure only (see sais.c).
 *
 * zlib's Adler-32 (RFC 1950 §8.2), which libbsc's adler32.cpp computes for
 * the block header words +16/+20/+24 (libbsc absent from the reference
 * tree; words +20 and +24 of the libbsc archive under tests/golden equal
 * zlib.adler32 of the payload and of header bytes 0..23).
 *
 * Pinned by: Python's zlib.adler32 on random inputs (tests/test_oracle.py).
 */
#include <stdint.h>

uint32_t orc_adler32(const uint8_t *p, int64_t n)
{
    uint32_t a = 1, b = 0;
    while (n > 0) {
        int64_t k = n < 5552 ? n : 5552;
        n -= k;
        while (k--) { a += *p++; b += a; }
        a %= 65521u;
        b %= 65521u;
    }
    return a | (b << 16);
}
