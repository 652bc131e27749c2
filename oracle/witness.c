/*
 * BLAKE3 R1CS witness generation — CPU oracle restatement (TEST
 * INFRASTRUCTURE ONLY; only tests, the smoke check and the benchmark's
 * CPU-baseline leg load it).
 *
 * Reference: CP/r1cs_hashes/blake3.rs (CP = crates/flock-prover/src/) and
 * CP/r1cs_hashes/common.rs.  One compression occupies a block of K = 2^14
 * witness bits (blake3.rs:111-113); the bit positions are the I/O-aligned
 * layout of blake3.rs:189-241:
 *
 *   [0, 256)        cv, 8 words                (input rows)
 *   [256, 512)      out_lo = state[w] ^ state[w+8]   (lin rows)
 *   512             the constant wire, 1
 *   [513, 1025)     m, 16 words                (input rows)
 *   1025, 1057, 1089, 1121   counter_lo, counter_hi, block_len, flags
 *   1153 + 250 g    G call g (g < 56): six 31-bit carry_aux records
 *                   (tmp0, a1, c1, tmp1, a2, c2), then the lin words b_new, d_new
 *   [15153, 15409)  out_hi = state[w+8] ^ cv[w]  (lin rows)
 *   [15409, 16384)  padding rows, all zero
 *
 * Row witnesses (blake3.rs:1886-1893; C = I so c == z): an input or lin row
 * holds (z, a, b) = (v, v, 1); the constant row (1, 1, 1); the carry row i of
 * an addition x + y holds (carry_aux_i, x_i ^ cin_i, y_i ^ cin_i) with
 * cin_i = XOR_{j<i} carry_aux_j, i.e. the word form of common.rs:80-88:
 * cin = (x + y) ^ x ^ y, left = (x ^ cin) & 0x7fffffff, right likewise,
 * carry_aux = left & right.  Bits are packed LSB-first, bit p of a block at
 * u64 word p / 64, bit p % 64 (the F128 packing of blake3.rs:2044-2182).
 *
 * Batch driver (common.rs:196-324 with the padding of blake3.rs:4617-4647):
 * 2^n_blocks_log slots (>= 8); slots past n_blocks hold the witness of the
 * all-zero compression (cv = 0, m = 0, counter 0, block_len 0, flags 0).  The
 * lincheck stripe is the 8-block bit transpose of z (common.rs:304-320 with
 * the byte order of CF/bits.rs:22-67 and CF/zerocheck/univariate_skip_optimized
 * /kernels/portable.rs:12-23): stripe byte g * K + p has bit r = z bit p of
 * block 8 g + r.
 *
 * This file is written for clarity: bits are placed one at a time.
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

#define WK_LOG 14
#define WK (1u << WK_LOG)
#define WORDS_PER_BLOCK (WK / 64)

enum {
    W_CV_BASE = 0,
    W_OUT_LO_BASE = 256,
    W_CONST_POS = 512,
    W_M_BASE = 513,
    W_T_LO_BASE = 1025,
    W_T_HI_BASE = 1057,
    W_BLEN_BASE = 1089,
    W_FLAGS_BASE = 1121,
    W_GS_BASE = 1153,
    W_G_STRIDE = 250,
    W_N_G = 56,
    W_OUT_HI_BASE = 1153 + 56 * 250,
    W_USEFUL_BITS = 1153 + 56 * 250 + 256
};

static const uint32_t B3_IV[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                                  0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
static const int B3_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
/* G lanes (a, b, c, d): four column then four diagonal calls (blake3.rs:153-162). */
static const int G_LANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                                  {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

void orc_witness_layout(size_t out[6])
{
    out[0] = WK;
    out[1] = W_CONST_POS;
    out[2] = W_GS_BASE;
    out[3] = W_G_STRIDE;
    out[4] = W_OUT_HI_BASE;
    out[5] = W_USEFUL_BITS;
}

static void set_bit(uint64_t *v, unsigned pos, unsigned bit)
{
    if (bit)
        v[pos >> 6] |= 1ull << (pos & 63);
}

/* Row interval [pos, pos + width) gets (z, a, b) = (zv, av, bv) bitwise. */
static void put_row(uint64_t *z, uint64_t *a, uint64_t *b, unsigned pos, unsigned width,
                    uint32_t zv, uint32_t av, uint32_t bv)
{
    for (unsigned i = 0; i < width; ++i) {
        set_bit(z, pos + i, (zv >> i) & 1u);
        set_bit(a, pos + i, (av >> i) & 1u);
        set_bit(b, pos + i, (bv >> i) & 1u);
    }
}

static void put_lin(uint64_t *z, uint64_t *a, uint64_t *b, unsigned pos, uint32_t v)
{
    put_row(z, a, b, pos, 32, v, v, 0xffffffffu);
}

static uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

/* One 32-bit addition with its carry records at `pos`, bit by bit: the carry
 * chain of blake3.rs:1739-1753 (carry_aux_i = (x_i ^ c_i)(y_i ^ c_i),
 * c_{i+1} = carry_aux_i ^ c_i) for i < 31; returns x + y mod 2^32. */
static uint32_t put_add(uint64_t *z, uint64_t *a, uint64_t *b, unsigned pos, uint32_t x, uint32_t y)
{
    uint32_t c = 0, sum = 0;
    for (unsigned i = 0; i < 32; ++i) {
        uint32_t xi = (x >> i) & 1u, yi = (y >> i) & 1u;
        sum |= (xi ^ yi ^ c) << i;
        uint32_t l = xi ^ c, r = yi ^ c, aux = l & r;
        if (i < 31) {
            set_bit(z, pos + i, aux);
            set_bit(a, pos + i, l);
            set_bit(b, pos + i, r);
        }
        c = aux ^ c;
    }
    return sum;
}

void orc_blake3_block_witness(const orc_compression *blk, uint64_t *z, uint64_t *a, uint64_t *b)
{
    memset(z, 0, WORDS_PER_BLOCK * 8);
    memset(a, 0, WORDS_PER_BLOCK * 8);
    memset(b, 0, WORDS_PER_BLOCK * 8);
    const uint32_t t_lo = (uint32_t)blk->counter, t_hi = (uint32_t)(blk->counter >> 32);

    for (int w = 0; w < 8; ++w)
        put_lin(z, a, b, W_CV_BASE + 32 * w, blk->cv[w]);
    put_row(z, a, b, W_CONST_POS, 1, 1, 1, 1);
    for (int i = 0; i < 16; ++i)
        put_lin(z, a, b, W_M_BASE + 32 * i, blk->m[i]);
    put_lin(z, a, b, W_T_LO_BASE, t_lo);
    put_lin(z, a, b, W_T_HI_BASE, t_hi);
    put_lin(z, a, b, W_BLEN_BASE, blk->block_len);
    put_lin(z, a, b, W_FLAGS_BASE, blk->flags);

    uint32_t s[16];
    for (int w = 0; w < 8; ++w)
        s[w] = blk->cv[w];
    for (int i = 0; i < 4; ++i)
        s[8 + i] = B3_IV[i];
    s[12] = t_lo;
    s[13] = t_hi;
    s[14] = blk->block_len;
    s[15] = blk->flags;
    uint32_t msg[16];
    memcpy(msg, blk->m, sizeof msg);

    for (int r = 0; r < 7; ++r) {
        for (int gi = 0; gi < 8; ++gi) {
            const unsigned base = W_GS_BASE + W_G_STRIDE * (unsigned)(8 * r + gi);
            const int la = G_LANES[gi][0], lb = G_LANES[gi][1], lc = G_LANES[gi][2], ld = G_LANES[gi][3];
            const uint32_t mx = msg[2 * gi], my = msg[2 * gi + 1];
            uint32_t tmp0 = put_add(z, a, b, base + 0 * 31, s[la], s[lb]);
            uint32_t a1 = put_add(z, a, b, base + 1 * 31, tmp0, mx);
            uint32_t d1 = rotr(s[ld] ^ a1, 16);
            uint32_t c1 = put_add(z, a, b, base + 2 * 31, s[lc], d1);
            uint32_t b1 = rotr(s[lb] ^ c1, 12);
            uint32_t tmp1 = put_add(z, a, b, base + 3 * 31, a1, b1);
            uint32_t a2 = put_add(z, a, b, base + 4 * 31, tmp1, my);
            uint32_t d2 = rotr(d1 ^ a2, 8);
            uint32_t c2 = put_add(z, a, b, base + 5 * 31, c1, d2);
            uint32_t bn = rotr(b1 ^ c2, 7);
            put_lin(z, a, b, base + 186, bn);
            put_lin(z, a, b, base + 218, d2);
            s[la] = a2;
            s[lb] = bn;
            s[lc] = c2;
            s[ld] = d2;
        }
        uint32_t nm[16];
        for (int i = 0; i < 16; ++i)
            nm[i] = msg[B3_PERM[i]];
        memcpy(msg, nm, sizeof msg);
    }
    for (int w = 0; w < 8; ++w) {
        put_lin(z, a, b, W_OUT_LO_BASE + 32 * w, s[w] ^ s[w + 8]);
        put_lin(z, a, b, W_OUT_HI_BASE + 32 * w, s[w + 8] ^ blk->cv[w]);
    }
}

int orc_blake3_witness(const orc_compression *blocks, size_t n_blocks, size_t n_blocks_log,
                       uint64_t *z, uint64_t *a, uint64_t *b, uint8_t *stripe)
{
    if (n_blocks_log < 3 || n_blocks_log > 40)
        return 1;
    const size_t n_total = (size_t)1 << n_blocks_log;
    if (n_blocks > n_total)
        return 1;
    orc_compression pad;
    memset(&pad, 0, sizeof pad);
    for (size_t j = 0; j < n_total; ++j) {
        const orc_compression *blk = j < n_blocks ? &blocks[j] : &pad;
        orc_blake3_block_witness(blk, z + j * WORDS_PER_BLOCK, a + j * WORDS_PER_BLOCK,
                                 b + j * WORDS_PER_BLOCK);
    }
    if (stripe) {
        for (size_t g = 0; g < n_total / 8; ++g) {
            for (unsigned p = 0; p < WK; ++p) {
                uint8_t v = 0;
                for (unsigned r = 0; r < 8; ++r) {
                    const uint64_t *zb = z + (8 * g + r) * WORDS_PER_BLOCK;
                    v |= (uint8_t)(((zb[p >> 6] >> (p & 63)) & 1u) << r);
                }
                stripe[g * WK + p] = v;
            }
        }
    }
    return 0;
}
