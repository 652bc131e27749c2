/*
 * Ligerito opening glue — CPU oracle restatement (TEST INFRASTRUCTURE ONLY;
 * only tests, the smoke check and the benchmark's CPU-baseline leg load it).
 *
 * Reference (CF = crates/flock-core/src/):
 *   merkle_multi_proof / _sibling_indices   CF/merkle.rs:1020-1107: positions
 *       sorted and deduplicated; level by level, a position whose sibling is
 *       not active emits the sibling's flat index; parents p >> 1 form the
 *       next level's active set.
 *   verify_merkle_multi_proof               CF/merkle.rs:1114-1191
 *   query row gather                        the opened rows of the queried
 *       positions: row p = codeword[p * row_len .. (p + 1) * row_len)
 *       (SoA layout, CF/ntt/additive_ntt_f128.rs:2700-2720).
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static int cmp_u64(const void *a, const void *b)
{
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

/* Sorted unique copy of positions into act (capacity n); returns the count. */
static size_t sort_unique(const uint64_t *pos, size_t n, uint64_t *act)
{
    memcpy(act, pos, n * sizeof *act);
    qsort(act, n, sizeof *act, cmp_u64);
    size_t m = 0;
    for (size_t i = 0; i < n; i++)
        if (m == 0 || act[m - 1] != act[i])
            act[m++] = act[i];
    return m;
}

size_t orc_merkle_multi_proof_indices(size_t num_leaves, const uint64_t *pos, size_t n, uint64_t *out)
{
    if (n == 0 || num_leaves <= 1)
        return 0;
    uint64_t *act = malloc(n * sizeof *act);
    size_t m = sort_unique(pos, n, act), k = 0, level_start = 0, level_len = num_leaves;
    while (level_len > 1) {
        size_t w = 0;
        for (size_t i = 0; i < m;) {
            uint64_t p = act[i];
            if (i + 1 < m && act[i + 1] == (p ^ 1)) {
                i += 2;
            } else {
                out[k++] = level_start + (p ^ 1);
                i += 1;
            }
            act[w++] = p >> 1;
        }
        m = w;
        level_start += level_len;
        level_len >>= 1;
    }
    free(act);
    return k;
}

int orc_merkle_multi_verify(const uint8_t root[32], size_t num_leaves, const uint64_t *pos, const uint8_t *leaf_hashes,
                            size_t n, const uint8_t *proof, size_t n_proof)
{
    if (!num_leaves || (num_leaves & (num_leaves - 1)))
        return 0;
    if (n == 0)
        return n_proof == 0;
    for (size_t i = 0; i < n; i++)
        if (pos[i] >= num_leaves || (i && pos[i - 1] >= pos[i]))
            return 0;
    if (num_leaves == 1)
        return n_proof == 0 && !memcmp(leaf_hashes, root, 32);
    uint64_t *p = malloc(n * sizeof *p);
    uint8_t *h = malloc(n * 32);
    memcpy(p, pos, n * sizeof *p);
    memcpy(h, leaf_hashes, n * 32);
    size_t m = n, used = 0;
    int ok = 1;
    for (size_t level_len = num_leaves; level_len > 1 && ok; level_len >>= 1) {
        size_t w = 0;
        for (size_t i = 0; i < m && ok;) {
            uint8_t l[32], r[32];
            uint64_t q = p[i];
            if (i + 1 < m && p[i + 1] == (q ^ 1)) {
                memcpy(l, h + 32 * i, 32);
                memcpy(r, h + 32 * (i + 1), 32);
                i += 2;
            } else {
                if (used == n_proof) {
                    ok = 0;
                    break;
                }
                const uint8_t *sib = proof + 32 * used++;
                if (q & 1) {
                    memcpy(l, sib, 32);
                    memcpy(r, h + 32 * i, 32);
                } else {
                    memcpy(l, h + 32 * i, 32);
                    memcpy(r, sib, 32);
                }
                i += 1;
            }
            orc_blake3_parent_cv(l, r, h + 32 * w);
            p[w++] = q >> 1;
        }
        m = w;
    }
    ok = ok && used == n_proof && m == 1 && !memcmp(h, root, 32);
    free(p);
    free(h);
    return ok;
}

void orc_gather_rows(const orc_f128 *codeword, size_t row_len, const uint64_t *pos, size_t n, orc_f128 *out)
{
    for (size_t i = 0; i < n; i++)
        memcpy(out + i * row_len, codeword + pos[i] * row_len, row_len * sizeof *out);
}
