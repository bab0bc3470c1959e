/*
 * ORACLE — test infrastructure only. Nothing in the product links or calls
 * this file; tests, the smoke check and bench.py's CPU-baseline leg use it as
 * the checker.
 *
 * Suffix array by induced sorting (SA-IS: Nong, Zhang & Chan, "Two Efficient
 * Algorithms for Linear Time Suffix Array Construction", IEEE TC 2011).
 *
 * Convention: a virtual sentinel, smaller than every symbol, follows T; the
 * sentinel's own (empty) suffix is NOT stored, so SA has exactly n entries and
 * a suffix that is a proper prefix of another sorts first. This is the order
 * libsais (the suffix sorter libbsc's forward BWT calls from
 * libbsc/bwt/bwt.cpp, absent from the reference tree; referenced by
 * crates/qz-lib/build.rs:25) produces.
 *
 * Pinned by: brute-force suffix sorting on random and adversarial small inputs
 * (tests/test_oracle.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CH(i) (cs == 1 ? (int32_t)((const uint8_t *)T)[i] : ((const int32_t *)T)[i])

static void buckets(const void *T, int cs, int32_t n, int32_t K, int32_t *bkt, int end)
{
    memset(bkt, 0, (size_t)K * sizeof(int32_t));
    for (int32_t i = 0; i < n; ++i) bkt[CH(i)]++;
    int32_t sum = 0;
    for (int32_t c = 0; c < K; ++c) {
        sum += bkt[c];
        bkt[c] = end ? sum : sum - bkt[c];
    }
}

static inline int is_lms(const uint8_t *t, int32_t i) { return i > 0 && t[i] && !t[i - 1]; }

/* t[i] = 1 for S-type, 0 for L-type. */
static void induce(const void *T, int cs, int32_t *SA, const uint8_t *t, int32_t *bkt, int32_t n, int32_t K)
{
    buckets(T, cs, n, K, bkt, 0);
    /* the suffix in front of the virtual sentinel is L-type and comes first */
    SA[bkt[CH(n - 1)]++] = n - 1;
    for (int32_t i = 0; i < n; ++i) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && !t[j]) SA[bkt[CH(j)]++] = j;
    }
    buckets(T, cs, n, K, bkt, 1);
    for (int32_t i = n - 1; i >= 0; --i) {
        int32_t j = SA[i] - 1;
        if (SA[i] > 0 && t[j]) SA[--bkt[CH(j)]] = j;
    }
}

static int sais_rec(const void *T, int cs, int32_t *SA, int32_t n, int32_t K)
{
    if (n == 1) { SA[0] = 0; return 0; }
    uint8_t *t = (uint8_t *)malloc((size_t)n);
    int32_t *bkt = (int32_t *)malloc((size_t)K * sizeof(int32_t));
    if (!t || !bkt) { free(t); free(bkt); return -2; }

    t[n - 1] = 0;
    for (int32_t i = n - 2; i >= 0; --i) {
        int32_t a = CH(i), b = CH(i + 1);
        t[i] = (uint8_t)(a < b || (a == b && t[i + 1]));
    }

    /* stage 1: sort LMS substrings */
    buckets(T, cs, n, K, bkt, 1);
    for (int32_t i = 0; i < n; ++i) SA[i] = -1;
    for (int32_t i = 1; i < n; ++i)
        if (is_lms(t, i)) SA[--bkt[CH(i)]] = i;
    induce(T, cs, SA, t, bkt, n, K);

    int32_t n1 = 0;
    for (int32_t i = 0; i < n; ++i)
        if (is_lms(t, SA[i])) SA[n1++] = SA[i];

    /* name the LMS substrings */
    for (int32_t i = n1; i < n; ++i) SA[i] = -1;
    int32_t name = 0, prev = -1;
    for (int32_t i = 0; i < n1; ++i) {
        int32_t pos = SA[i], diff = 0;
        for (int32_t d = 0;; ++d) {
            if (prev == -1 || pos + d == n || prev + d == n || CH(pos + d) != CH(prev + d) ||
                t[pos + d] != t[prev + d]) {
                diff = 1;
                break;
            }
            if (d > 0 && (is_lms(t, pos + d) || is_lms(t, prev + d))) break;
        }
        if (diff) { name++; prev = pos; }
        SA[n1 + pos / 2] = name - 1;
    }
    for (int32_t i = n - 1, j = n - 1; i >= n1; --i)
        if (SA[i] >= 0) SA[j--] = SA[i];

    /* stage 2: order the LMS suffixes */
    int32_t *s1 = SA + n - n1;
    if (name < n1) {
        int rc = sais_rec(s1, 4, SA, n1, name);
        if (rc) { free(t); free(bkt); return rc; }
    } else {
        for (int32_t i = 0; i < n1; ++i) SA[s1[i]] = i;
    }

    /* stage 3: induce the full order from the sorted LMS suffixes */
    buckets(T, cs, n, K, bkt, 1);
    for (int32_t i = 1, j = 0; i < n; ++i)
        if (is_lms(t, i)) s1[j++] = i;
    for (int32_t i = 0; i < n1; ++i) SA[i] = s1[SA[i]];
    for (int32_t i = n1; i < n; ++i) SA[i] = -1;
    for (int32_t i = n1 - 1; i >= 0; --i) {
        int32_t j = SA[i];
        SA[i] = -1;
        SA[--bkt[CH(j)]] = j;
    }
    induce(T, cs, SA, t, bkt, n, K);

    free(t);
    free(bkt);
    return 0;
}

int orc_suffix_array(const uint8_t *T, int32_t *SA, int32_t n)
{
    if (n < 0) return -1;
    if (n == 0) return 0;
    return sais_rec(T, 1, SA, n, 256);
}
