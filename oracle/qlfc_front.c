/*
 * ORACLE — test infrastructure only (see sais.c).
 *
 * QLFC front end: the run/rank transform libbsc's QLFC coders apply to each
 * coder segment of the BWT output (libbsc coder/qlfc/qlfc.cpp, absent from
 * the reference tree; compiled by crates/qz-lib/build.rs:19-37).
 *
 * The segment is cut into maximal runs. Symbols are kept in a list ordered by
 * their NEXT occurrence. Scanning runs from last to first, the rank of a run
 * of symbol c is c's position in that list (0 for the front), or the list
 * length when c does not occur again; c then moves to the front. The list
 * left after the scan (symbols by first occurrence) is the initial table the
 * decoder starts from: the decoder's current symbol is always the table
 * front, and after each run it moves that symbol back to position `rank`.
 *
 * Pinned by: the closed form (decode(encode(x)) == x on random and adversarial
 * inputs), the two direct characterisations checked in tests (rank ==
 * #distinct symbols strictly between the run and the symbol's next
 * occurrence), and the table of the libbsc archive kept under tests/golden,
 * whose first coder segment opens with this table (C A T G N, terminated by
 * repeating N).
 */
#include <stdint.h>
#include <string.h>

/*
 * Forward transform. Writes per run (forward order) its symbol, rank and
 * length. table[0..nsym-1] receives the initial list; returns the run count,
 * *nsym the number of distinct symbols.
 */
int32_t orc_qlfc_transform(const uint8_t *in, int32_t n, uint8_t *run_sym, uint8_t *run_rank,
                           int32_t *run_len, uint8_t *table, int32_t *nsym)
{
    int32_t runs = 0;
    for (int32_t i = 0; i < n;) {
        int32_t j = i + 1;
        while (j < n && in[j] == in[i]) ++j;
        run_sym[runs] = in[i];
        run_len[runs] = j - i;
        ++runs;
        i = j;
    }
    uint8_t list[256];
    int32_t len = 0;
    for (int32_t t = runs - 1; t >= 0; --t) {
        uint8_t c = run_sym[t];
        int32_t pos = 0;
        while (pos < len && list[pos] != c) ++pos;
        run_rank[t] = (uint8_t)pos;
        if (pos == len) ++len;
        memmove(list + 1, list, (size_t)pos);
        list[0] = c;
    }
    memcpy(table, list, (size_t)len);
    *nsym = len;
    return runs;
}

/* Inverse transform: rebuild the segment from table, ranks and lengths. */
int32_t orc_qlfc_inverse(const uint8_t *table, int32_t nsym, const uint8_t *run_rank, const int32_t *run_len,
                         int32_t runs, uint8_t *out, int32_t n)
{
    uint8_t list[256];
    memcpy(list, table, (size_t)nsym);
    int32_t pos = 0;
    for (int32_t t = 0; t < runs; ++t) {
        uint8_t c = list[0];
        int32_t r = run_rank[t];
        if (r >= nsym && !(r == 0)) return -6;
        if (pos + run_len[t] > n) return -6;
        memset(out + pos, c, (size_t)run_len[t]);
        pos += run_len[t];
        if (r > 0) {
            memmove(list, list + 1, (size_t)r);
            list[r] = c;
        }
    }
    return pos == n ? 0 : -6;
}

/*
 * Coder segments of one block (libbsc coder.cpp, absent from the reference
 * tree). Count: one segment up to 512 KiB, then one per started 512 KiB, at
 * most 8 — fitted to the three blocks of the libbsc archive under
 * tests/golden (2,904,234 B -> 6; 4,450,000 and 5,050,000 B -> 8) and
 * "parity unpinned" elsewhere. Boundaries: `changes` = number of
 * i = 1, 33, 65, ... (< n) with in[i] != in[i-1]; if changes > nb, boundary
 * m (1..nb-1) sits at the sample of the (m * (changes / nb))-th change;
 * otherwise the block splits evenly. Pinned only by the archive's segment
 * sizes being consistent with it (every boundary is 1 mod 32).
 */
int32_t orc_num_segments(int32_t n)
{
    int64_t nb = ((int64_t)n + 524287) >> 19;
    return nb < 1 ? 1 : (nb > 8 ? 8 : (int32_t)nb);
}

void orc_split_segments(const uint8_t *in, int32_t n, int32_t nb, int32_t *start, int32_t *size)
{
    int32_t changes = 0;
    for (int32_t i = 1; i < n; i += 32)
        if (in[i] != in[i - 1]) ++changes;
    if (nb > 1 && changes > nb) {
        int32_t per = changes / nb, seen = 0, id = 0;
        start[0] = 0;
        for (int32_t i = 1; i < n && id < nb - 1; i += 32) {
            if (in[i] != in[i - 1] && ++seen % per == 0) {
                size[id] = i - start[id];
                start[++id] = i;
            }
        }
        size[nb - 1] = n - start[nb - 1];
    } else {
        for (int32_t p = 0; p < nb; ++p) {
            start[p] = (n / nb) * p;
            size[p] = (p != nb - 1) ? n / nb : n - (n / nb) * (nb - 1);
        }
    }
}
