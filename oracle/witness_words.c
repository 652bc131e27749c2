/*
 * BLAKE3 R1CS witness, word form — CPU oracle (TEST INFRASTRUCTURE ONLY; only
 * tests, the smoke check and the benchmark's CPU-baseline leg load it).
 *
 * The same rows as witness.c, produced the way the reference's packed builder
 * does (CP/r1cs_hashes/blake3.rs:1919-2010,2044-2182, CP =
 * crates/flock-prover/src/): whole 32-bit row records appended to u64 words
 * in position order, carry records from the word identities of
 * CP/r1cs_hashes/common.rs:80-88.  witness.c places every bit on its own; the
 * tests hold the two to each other.  orc_blake3_witness_mt splits the slots
 * over POSIX threads (the benchmark's CPU baseline).
 */
#include <pthread.h>
#include <string.h>

#include "flock_oracle.h"

typedef struct {
    uint64_t *z, *a, *b;
    uint64_t cz, ca, cb;  /* pending low bits of the current word */
    unsigned used, widx;
} rows_t;

static void rows_push(rows_t *r, unsigned width, uint32_t zv, uint32_t av, uint32_t bv)
{
    r->cz |= (uint64_t)zv << r->used;
    r->ca |= (uint64_t)av << r->used;
    r->cb |= (uint64_t)bv << r->used;
    r->used += width;
    if (r->used >= 64) {
        r->z[r->widx] = r->cz;
        r->a[r->widx] = r->ca;
        r->b[r->widx] = r->cb;
        r->widx++;
        r->used -= 64;
        unsigned sh = width - r->used;
        r->cz = (uint64_t)zv >> sh;
        r->ca = (uint64_t)av >> sh;
        r->cb = (uint64_t)bv >> sh;
    }
}

static uint32_t rows_add(rows_t *r, uint32_t x, uint32_t y)
{
    uint32_t sum = x + y, cin = sum ^ x ^ y;
    uint32_t left = (x ^ cin) & 0x7fffffffu, right = (y ^ cin) & 0x7fffffffu;
    rows_push(r, 31, left & right, left, right);
    return sum;
}

static void rows_lin(rows_t *r, uint32_t v) { rows_push(r, 32, v, v, 0xffffffffu); }

static uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static const uint32_t WIV[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
static const int WPERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
static const int WLANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                 {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

void orc_blake3_block_witness_words(const orc_compression *blk, uint64_t *z, uint64_t *a, uint64_t *b)
{
    const uint32_t t_lo = (uint32_t)blk->counter, t_hi = (uint32_t)(blk->counter >> 32);
    uint32_t s[16], m[16], e[16];
    /* the compression first: out_lo (words 4..7) precedes the G records */
    memcpy(s, blk->cv, 32);
    memcpy(s + 8, WIV, 16);
    s[12] = t_lo, s[13] = t_hi, s[14] = blk->block_len, s[15] = blk->flags;
    memcpy(e, s, sizeof e);
    memcpy(m, blk->m, sizeof m);
    for (int r = 0; r < 7; ++r) {
        for (int g = 0; g < 8; ++g) {
            const int la = WLANES[g][0], lb = WLANES[g][1], lc = WLANES[g][2], ld = WLANES[g][3];
            e[la] = e[la] + e[lb] + m[2 * g];
            e[ld] = rotr32(e[ld] ^ e[la], 16);
            e[lc] = e[lc] + e[ld];
            e[lb] = rotr32(e[lb] ^ e[lc], 12);
            e[la] = e[la] + e[lb] + m[2 * g + 1];
            e[ld] = rotr32(e[ld] ^ e[la], 8);
            e[lc] = e[lc] + e[ld];
            e[lb] = rotr32(e[lb] ^ e[lc], 7);
        }
        uint32_t p[16];
        for (int i = 0; i < 16; ++i) p[i] = m[WPERM[i]];
        memcpy(m, p, sizeof m);
    }
    rows_t r = {z, a, b, 0, 0, 0, 0, 0};
    for (int i = 0; i < 4; ++i) rows_push(&r, 32, blk->cv[2 * i], blk->cv[2 * i], ~0u), rows_push(&r, 32, blk->cv[2 * i + 1], blk->cv[2 * i + 1], ~0u);
    for (int i = 0; i < 8; ++i) rows_lin(&r, e[i] ^ e[i + 8]);
    rows_push(&r, 1, 1, 1, 1);                         /* the constant wire */
    for (int i = 0; i < 16; ++i) rows_lin(&r, blk->m[i]);
    rows_lin(&r, t_lo);
    rows_lin(&r, t_hi);
    rows_lin(&r, blk->block_len);
    rows_lin(&r, blk->flags);
    memcpy(m, blk->m, sizeof m);
    for (int rd = 0; rd < 7; ++rd) {
        for (int g = 0; g < 8; ++g) {
            const int la = WLANES[g][0], lb = WLANES[g][1], lc = WLANES[g][2], ld = WLANES[g][3];
            uint32_t tmp0 = rows_add(&r, s[la], s[lb]);
            uint32_t a1 = rows_add(&r, tmp0, m[2 * g]);
            uint32_t d1 = rotr32(s[ld] ^ a1, 16);
            uint32_t c1 = rows_add(&r, s[lc], d1);
            uint32_t b1 = rotr32(s[lb] ^ c1, 12);
            uint32_t tmp1 = rows_add(&r, a1, b1);
            uint32_t a2 = rows_add(&r, tmp1, m[2 * g + 1]);
            uint32_t d2 = rotr32(d1 ^ a2, 8);
            uint32_t c2 = rows_add(&r, c1, d2);
            uint32_t bn = rotr32(b1 ^ c2, 7);
            rows_lin(&r, bn);
            rows_lin(&r, d2);
            s[la] = a2, s[lb] = bn, s[lc] = c2, s[ld] = d2;
        }
        uint32_t p[16];
        for (int i = 0; i < 16; ++i) p[i] = m[WPERM[i]];
        memcpy(m, p, sizeof m);
    }
    for (int w = 0; w < 8; ++w) rows_lin(&r, s[w + 8] ^ blk->cv[w]);
    /* the last partial word, then the zero padding rows */
    z[r.widx] = r.cz, a[r.widx] = r.ca, b[r.widx] = r.cb;
    for (unsigned w = r.widx + 1; w < 256; ++w) z[w] = a[w] = b[w] = 0;
}

typedef struct {
    const orc_compression *blocks;
    size_t n_blocks, lo, hi;
    uint64_t *z, *a, *b;
    uint8_t *stripe;
} wjob_t;

static void *wjob_main(void *p)
{
    wjob_t *j = p;
    orc_compression pad;
    memset(&pad, 0, sizeof pad);
    for (size_t s = j->lo; s < j->hi; ++s)
        orc_blake3_block_witness_words(s < j->n_blocks ? &j->blocks[s] : &pad, j->z + 256 * s, j->a + 256 * s,
                                       j->b + 256 * s);
    if (j->stripe) {
        /* byte c of the 8 words -> an 8x8 bit matrix (row r = block 8g + r),
         * transposed (Hacker's Delight 7-3): byte t = stripe byte 64 w + 8 c + t */
        for (size_t g = j->lo / 8; g < j->hi / 8; ++g)
            for (unsigned w = 0; w < 256; ++w)
                for (unsigned c = 0; c < 8; ++c) {
                    uint64_t x = 0;
                    for (unsigned r = 0; r < 8; ++r)
                        x |= ((j->z[256 * (8 * g + r) + w] >> (8 * c)) & 0xffu) << (8 * r);
                    uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
                    x = x ^ t ^ (t << 7);
                    t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
                    x = x ^ t ^ (t << 14);
                    t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
                    x = x ^ t ^ (t << 28);
                    for (unsigned b = 0; b < 8; ++b)
                        j->stripe[g * 16384 + 64 * w + 8 * c + b] = (uint8_t)(x >> (8 * b));
                }
    }
    return NULL;
}

int orc_blake3_witness_mt(const orc_compression *blocks, size_t n_blocks, size_t n_blocks_log, uint64_t *z,
                          uint64_t *a, uint64_t *b, uint8_t *stripe, size_t n_threads)
{
    if (n_blocks_log < 3 || n_blocks_log > 40)
        return 1;
    const size_t n_total = (size_t)1 << n_blocks_log;
    if (n_blocks > n_total)
        return 1;
    if (n_threads < 1)
        n_threads = 1;
    if (n_threads > 256)
        n_threads = 256;
    const size_t groups = n_total / 8;
    if (n_threads > groups)
        n_threads = groups;
    wjob_t jobs[256];
    pthread_t th[256];
    for (size_t t = 0; t < n_threads; ++t) {
        jobs[t] = (wjob_t){blocks, n_blocks, 8 * (groups * t / n_threads), 8 * (groups * (t + 1) / n_threads),
                           z, a, b, stripe};
    }
    size_t started = 0;
    for (size_t t = 1; t < n_threads; ++t, ++started)
        if (pthread_create(&th[t], NULL, wjob_main, &jobs[t]) != 0)
            break;
    wjob_main(&jobs[0]);
    for (size_t t = started + 1; t < n_threads; ++t)
        wjob_main(&jobs[t]); /* threads that did not start run here */
    for (size_t t = 1; t <= started; ++t)
        pthread_join(th[t], NULL);
    return 0;
}
