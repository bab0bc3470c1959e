/*
 * ORACLE — test infrastructure only (see sais.c).
 *
 * Forward and inverse Burrows-Wheeler transform with the conventions libbsc
 * uses. libbsc's bwt.cpp (absent from the reference tree; compiled by
 * crates/qz-lib/build.rs:19-37) calls libsais' bwt_aux / unbwt:
 *
 *   - rows are the n+1 suffixes of T$ in sorted order ($ smallest); row 0 is
 *     the empty suffix, whose BWT symbol is T[n-1];
 *   - the output U has n bytes: the full n+1-row BWT with the $ removed;
 *   - primary index p = 1 + (0-based rank of suffix 0 among the n real
 *     suffixes) = row of suffix 0 in the n+1-row matrix, so 1 <= p <= n;
 *   - sampled indexes, with r the largest power of two <= n/8:
 *     I[k] = 1 + rank(k*r) for k = 0..(n-1)/r, I[0] = p.
 *
 * libbsc stores (n-1)/r sampled indexes in the block trailer; the trailer
 * layout and the sampling rule are pinned by the libbsc archive kept under
 * tests/golden (three blocks; trailer count == (n-1)/r on all three).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_suffix_array(const uint8_t *T, int32_t *SA, int32_t n);

/* Largest power of two <= n/8 (libbsc's sampling step); 1 when n < 16. */
int32_t orc_bwt_sample_rate(int32_t n)
{
    int32_t m = n / 8;
    m |= m >> 1; m |= m >> 2; m |= m >> 4; m |= m >> 8; m |= m >> 16;
    m >>= 1;
    return m + 1;
}

/* BWT from a suffix array. U may not alias T. Returns primary index. */
int32_t orc_bwt_from_sa(const uint8_t *T, const int32_t *SA, uint8_t *U, int32_t n, int32_t r, int32_t *I)
{
    if (n <= 0) return -1;
    int32_t p = -1;
    U[0] = T[n - 1];
    int32_t j = 1;
    for (int32_t k = 0; k < n; ++k) {
        int32_t s = SA[k];
        if (I && r > 0 && (s % r) == 0) I[s / r] = k + 1;
        if (s == 0) { p = k + 1; continue; }
        U[j++] = T[s - 1];
    }
    return p;
}

/* libsais_bwt_aux equivalent. Returns the primary index (>= 1) or < 0. */
int32_t orc_bwt_aux(const uint8_t *T, uint8_t *U, int32_t n, int32_t r, int32_t *I)
{
    if (n <= 0) return -1;
    int32_t *SA = (int32_t *)malloc((size_t)n * sizeof(int32_t));
    if (!SA) return -2;
    int rc = orc_suffix_array(T, SA, n);
    if (rc) { free(SA); return rc; }
    int32_t p = orc_bwt_from_sa(T, SA, U, n, r, I);
    free(SA);
    return p;
}

/*
 * libbsc bsc_bwt_encode (in place on T). Writes the (n-1)/r trailer indexes
 * and returns the primary index. The trailer entry for sample k (k >= 1) is
 * I[k] - 1 in libbsc's convention; the primary index is I[0].
 */
int32_t orc_bsc_bwt_encode(uint8_t *T, int32_t n, uint8_t *num_indexes, int32_t *indexes)
{
    int32_t r = orc_bwt_sample_rate(n);
    int32_t I[257];
    int32_t cnt = (n - 1) / r;
    if (cnt > 255) return -1;
    uint8_t *U = (uint8_t *)malloc((size_t)n);
    if (!U) return -2;
    int32_t p = orc_bwt_aux(T, U, n, r, I);
    if (p < 0) { free(U); return p; }
    memcpy(T, U, (size_t)n);
    free(U);
    *num_indexes = (uint8_t)cnt;
    for (int32_t t = 0; t < cnt; ++t) indexes[t] = I[t + 1] - 1;
    return p;
}

/*
 * Inverse BWT by LF-mapping (libsais_unbwt equivalent). U has n bytes and
 * primary index p (1..n). Writes T (may not alias U). Returns 0 or < 0.
 */
int32_t orc_unbwt(const uint8_t *U, uint8_t *T, int32_t n, int32_t p)
{
    if (n <= 0) return 0;
    if (p < 1 || p > n) return -1;
    /* full last column L[0..n], L[p] = $ */
    int32_t C[257];
    memset(C, 0, sizeof(C));
    for (int32_t i = 0; i < n; ++i) C[U[i] + 1]++;
    /* $ occupies row 0 of the first column: symbol c starts at 1 + #(< c) */
    int32_t sum = 1;
    for (int32_t c = 0; c < 256; ++c) { int32_t t = C[c + 1]; C[c + 1] = sum; sum += t; }
    int32_t *LF = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
    if (!LF) return -2;
    int32_t occ[256];
    for (int c = 0; c < 256; ++c) occ[c] = C[c + 1];
    for (int32_t row = 0; row <= n; ++row) {
        if (row == p) { LF[row] = 0; continue; }
        uint8_t c = U[row < p ? row : row - 1];
        LF[row] = occ[c]++;
    }
    /* walk from row 0 (the empty suffix): L[row] = T[pos-1] */
    int32_t row = 0;
    for (int32_t pos = n - 1; pos >= 0; --pos) {
        uint8_t c = U[row < p ? row : row - 1];
        T[pos] = c;
        row = LF[row];
    }
    free(LF);
    return row == p ? 0 : -6;
}
