/*
 * Oracle (test infrastructure): orc_zc_round1 (round1_naive,
 * CF/zerocheck/univariate_skip.rs:90-152) for k_skip = 6, computed by
 * n_threads POSIX threads for the CPU baseline of the benchmark.  The sums are
 * those of orc_zc_round1, regrouped:
 *   - the extension of the 64 bits of one x from S to Lambda (inverse NTT on S,
 *     forward NTT on Lambda) is GF(2)-linear in the bits, so it is the XOR of
 *     the columns M[., s] over the set bits s, where column s is the extension
 *     of the unit vector e_s; the columns are combined eight bits at a time
 *     from 256-entry tables;
 *   - sum_x eq(x) phi8(v_l(x)) = sum_v phi8(v) (sum over x with v_l(x) = v of
 *     eq(x)): each thread XORs eq(x) into a bucket per (l, v), and the buckets
 *     are multiplied by phi8(v) once at the end.
 * The oracle tests check it against orc_zc_round1 bit for bit.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

#define R1_ELL 64
#define R1_BUCKETS (2 * R1_ELL * 256)

struct r1_ctx {
    const uint8_t *a, *b, *c;
    size_t nx, n_threads;
    const orc_f128 *eq;
    const uint64_t (*tab)[256][R1_ELL / 8]; /* [8 chunks][byte value][64 bytes as words] */
    const uint8_t *mul;                      /* 256 x 256 GF(2^8) products */
    orc_f128 *buckets;                       /* n_threads x R1_BUCKETS */
};

static void extend(const uint64_t (*tab)[256][R1_ELL / 8], const uint8_t *bits, uint64_t out[R1_ELL / 8]) {
    memset(out, 0, R1_ELL);
    for (int j = 0; j < 8; j++) {
        const uint64_t *row = tab[j][bits[j]];
        for (int w = 0; w < R1_ELL / 8; w++) out[w] ^= row[w];
    }
}

static void r1_range(void *p, size_t t0, size_t t1) {
    struct r1_ctx *c = (struct r1_ctx *)p;
    for (size_t t = t0; t < t1; t++) {
        orc_f128 *bab = c->buckets + t * R1_BUCKETS, *bc = bab + R1_ELL * 256;
        memset(bab, 0, R1_BUCKETS * sizeof(orc_f128));
        size_t lo = t * c->nx / c->n_threads, hi = (t + 1) * c->nx / c->n_threads;
        for (size_t x = lo; x < hi; x++) {
            uint64_t ea[R1_ELL / 8], eb[R1_ELL / 8], ec[R1_ELL / 8];
            extend(c->tab, c->a + 8 * x, ea);
            extend(c->tab, c->b + 8 * x, eb);
            extend(c->tab, c->c + 8 * x, ec);
            const uint8_t *va = (const uint8_t *)ea, *vb = (const uint8_t *)eb, *vc = (const uint8_t *)ec;
            const orc_f128 e = c->eq[x];
            for (int i = 0; i < R1_ELL; i++) {
                orc_f128 *s = &bab[i * 256 + c->mul[(va[i] << 8) | vb[i]]];
                s->lo ^= e.lo;
                s->hi ^= e.hi;
                s = &bc[i * 256 + vc[i]];
                s->lo ^= e.lo;
                s->hi ^= e.hi;
            }
        }
    }
}

int orc_zc_round1_mt(const uint8_t *a, const uint8_t *b, const uint8_t *c, size_t m, size_t k_skip,
                     const orc_f128 *r, orc_f128 *p_ab, orc_f128 *p_c, size_t n_threads) {
    if (k_skip != 6 || m < k_skip) return -1;
    if (n_threads < 1) n_threads = 1;
    if (n_threads > 256) n_threads = 256;
    const size_t nx = (size_t)1 << (m - k_skip);
    if (n_threads > nx) n_threads = nx;
    uint64_t(*tab)[256][R1_ELL / 8] = malloc(8 * sizeof *tab);
    uint8_t *mul = malloc(65536);
    orc_f128 *eq = malloc(nx * sizeof(orc_f128));
    orc_f128 *buckets = malloc(n_threads * R1_BUCKETS * sizeof(orc_f128));
    int rc = -1;
    if (!tab || !mul || !eq || !buckets) goto out;
    /* column s of M: the extension of e_s */
    uint8_t col[R1_ELL][R1_ELL];
    for (int s = 0; s < R1_ELL; s++) {
        memset(col[s], 0, R1_ELL);
        col[s][s] = 1;
        if (orc_gf8_inverse(k_skip, 0, col[s]) || orc_gf8_forward(k_skip, R1_ELL, col[s])) goto out;
    }
    for (int j = 0; j < 8; j++)
        for (int v = 0; v < 256; v++) {
            uint8_t acc[R1_ELL] = {0};
            for (int bit = 0; bit < 8; bit++)
                if ((v >> bit) & 1)
                    for (int i = 0; i < R1_ELL; i++) acc[i] ^= col[8 * j + bit][i];
            memcpy(tab[j][v], acc, R1_ELL);
        }
    for (int x = 0; x < 256; x++)
        for (int y = 0; y < 256; y++) mul[(x << 8) | y] = orc_f8_mul((uint8_t)x, (uint8_t)y);
    orc_build_eq(r + k_skip, m - k_skip, eq);
    struct r1_ctx ctx = {a, b, c, nx, n_threads, eq, (const uint64_t(*)[256][R1_ELL / 8])tab, mul, buckets};
    orc_parallel_for(n_threads, n_threads, r1_range, &ctx);
    for (size_t t = 1; t < n_threads; t++)
        for (size_t i = 0; i < R1_BUCKETS; i++) {
            buckets[i].lo ^= buckets[t * R1_BUCKETS + i].lo;
            buckets[i].hi ^= buckets[t * R1_BUCKETS + i].hi;
        }
    orc_f128 phi[256];
    orc_phi8_table(phi);
    for (int i = 0; i < R1_ELL; i++) {
        orc_f128 sab = {0, 0}, sc = {0, 0};
        for (int v = 1; v < 256; v++) {
            orc_f128 t1 = orc_f128_mul(phi[v], buckets[i * 256 + v]);
            orc_f128 t2 = orc_f128_mul(phi[v], buckets[R1_ELL * 256 + i * 256 + v]);
            sab.lo ^= t1.lo;
            sab.hi ^= t1.hi;
            sc.lo ^= t2.lo;
            sc.hi ^= t2.hi;
        }
        p_ab[i] = sab;
        p_c[i] = sc;
    }
    rc = 0;
out:
    free(tab);
    free(mul);
    free(eq);
    free(buckets);
    return rc;
}
