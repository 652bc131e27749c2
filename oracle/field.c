/*
 * Oracle (test infrastructure): GF(2^128) GHASH-form arithmetic, GF(2^8) AES
 * arithmetic, the phi8 embedding and the SplitMix64 test RNG.
 *
 *   orc_clmul64            CF/field/gf2_128/portable.rs:4-20 (bit-serial carry-less product)
 *   orc_f128_mul_unreduced CF/field/gf2_128/portable.rs:22-35 (4-product schoolbook)
 *   orc_ghash_reduce       CF/field/gf2_128.rs:242-262 (fold r2:r3 by U^U<<1^U<<2^U<<7)
 *   orc_f128_mul_by_x      CF/field/gf2_128.rs:174-184
 *   orc_f128_inv           CF/field/gf2_128.rs:84-97 (value a^(2^128-2); 0 -> 0)
 *   orc_f8_mul / inv       CF/field/gf2_8.rs:20-120 (AES poly x^8+x^4+x^3+x+1)
 *   orc_phi8_table         CF/field/phi8.rs:19,1047 (embedding fixed by phi8(0x02))
 *   orc_splitmix64_next    CF/ntt.rs:175-187, benchmark-tools/common/src/lib.rs:4-18
 */
#include "flock_oracle.h"

/* Bit-serial carry-less product, exactly the reference's portable form. */
void orc_clmul64_bitwise(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
    uint64_t l = 0, h = 0;
    for (int i = 0; i < 64; i++) {
        if ((a >> i) & 1) {
            l ^= b << i;
            if (i) h ^= b >> (64 - i);
        }
    }
    *lo = l;
    *hi = h;
}

/* The same product with the x86 PCLMULQDQ instruction when the build enables it
 * (the reference's x86 path multiplies the same way, CF/field/gf2_128/x86_64.rs);
 * the bit-serial form above is the fallback and pins this one in the tests. */
#if defined(__PCLMUL__)
#include <wmmintrin.h>
void orc_clmul64(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) {
    __m128i p = _mm_clmulepi64_si128(_mm_set_epi64x(0, (long long)a), _mm_set_epi64x(0, (long long)b), 0x00);
    *lo = (uint64_t)_mm_cvtsi128_si64(p);
    *hi = (uint64_t)_mm_cvtsi128_si64(_mm_unpackhi_epi64(p, p));
}
#else
void orc_clmul64(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi) { orc_clmul64_bitwise(a, b, lo, hi); }
#endif

void orc_f128_mul_unreduced(orc_f128 a, orc_f128 b, uint64_t out[4]) {
    uint64_t ll_lo, ll_hi, lh_lo, lh_hi, hl_lo, hl_hi, hh_lo, hh_hi;
    orc_clmul64(a.lo, b.lo, &ll_lo, &ll_hi);
    orc_clmul64(a.lo, b.hi, &lh_lo, &lh_hi);
    orc_clmul64(a.hi, b.lo, &hl_lo, &hl_hi);
    orc_clmul64(a.hi, b.hi, &hh_lo, &hh_hi);
    out[0] = ll_lo;
    out[1] = ll_hi ^ lh_lo ^ hl_lo;
    out[2] = hh_lo ^ lh_hi ^ hl_hi;
    out[3] = hh_hi;
}

orc_f128 orc_ghash_reduce(const uint64_t r[4]) {
    uint64_t r2 = r[2], r3 = r[3];
    uint64_t t_lo = r2 ^ (r2 << 1) ^ (r2 << 2) ^ (r2 << 7);
    uint64_t t_hi = r3 ^ ((r3 << 1) | (r2 >> 63)) ^ ((r3 << 2) | (r2 >> 62)) ^ ((r3 << 7) | (r2 >> 57));
    uint64_t ov = (r3 >> 63) ^ (r3 >> 62) ^ (r3 >> 57);
    uint64_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
    orc_f128 o = {r[0] ^ t_lo ^ corr, r[1] ^ t_hi};
    return o;
}

orc_f128 orc_f128_mul(orc_f128 a, orc_f128 b) {
    uint64_t u[4];
    orc_f128_mul_unreduced(a, b, u);
    return orc_ghash_reduce(u);
}

void orc_f128_mul_many(const orc_f128 *a, const orc_f128 *b, orc_f128 *out, size_t n) {
    for (size_t i = 0; i < n; i++) out[i] = orc_f128_mul(a[i], b[i]);
}

orc_f128 orc_f128_square(orc_f128 a) { return orc_f128_mul(a, a); }

orc_f128 orc_f128_mul_by_x(orc_f128 z) {
    uint64_t carry = z.hi >> 63;
    uint64_t mask = 0 - carry;
    orc_f128 o = {(z.lo << 1) ^ (0x87 & mask), (z.hi << 1) | (z.lo >> 63)};
    return o;
}

orc_f128 orc_f128_inv(orc_f128 a) {
    /* a^(2^128 - 2) by square-and-multiply over the exponent bits
     * (bit 0 clear, bits 1..127 set). */
    orc_f128 result = {1, 0};
    orc_f128 base = a;
    for (int i = 0; i < 128; i++) {
        if (i >= 1) result = orc_f128_mul(result, base);
        base = orc_f128_square(base);
    }
    return result;
}

uint8_t orc_f8_mul(uint8_t a, uint8_t b) {
    uint16_t acc = 0;
    for (int i = 0; i < 8; i++)
        if ((a >> i) & 1) acc ^= (uint16_t)((uint16_t)b << i);
    /* reduce by x^8 + x^4 + x^3 + x + 1 (0x11B), high bits first */
    for (int i = 15; i >= 8; i--)
        if ((acc >> i) & 1) acc ^= (uint16_t)(0x11B << (i - 8));
    return (uint8_t)acc;
}

uint8_t orc_f8_inv(uint8_t a) {
    /* a^254 */
    uint8_t r = 1, sq = a;
    for (int i = 0; i < 8; i++) {
        if ((0xFE >> i) & 1) r = orc_f8_mul(r, sq);
        sq = orc_f8_mul(sq, sq);
    }
    return r;
}

/* Image of the GF(2^8) element x (= 0x02) under the embedding used by the
 * reference (CF/field/phi8.rs:28-31, entry [2] of PHI_8_TABLE).  phi8 is a
 * ring homomorphism, so phi8(sum b_i x^i) = sum b_i phi8(x)^i. */
static const orc_f128 PHI8_OF_X = {0x6b8330483c2e9849ull, 0x0dcb364640a222feull};

void orc_phi8_table(orc_f128 out[256]) {
    orc_f128 pw[8];
    pw[0].lo = 1;
    pw[0].hi = 0;
    for (int i = 1; i < 8; i++) pw[i] = orc_f128_mul(pw[i - 1], PHI8_OF_X);
    for (int v = 0; v < 256; v++) {
        orc_f128 acc = {0, 0};
        for (int i = 0; i < 8; i++)
            if ((v >> i) & 1) {
                acc.lo ^= pw[i].lo;
                acc.hi ^= pw[i].hi;
            }
        out[v] = acc;
    }
}

uint64_t orc_splitmix64_next(uint64_t *state) {
    *state += 0x9E3779B97F4A7C15ull;
    uint64_t z = *state;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void orc_fill_f128(uint64_t seed, orc_f128 *out, size_t n) {
    uint64_t s = seed;
    for (size_t i = 0; i < n; i++) {
        out[i].lo = orc_splitmix64_next(&s);
        out[i].hi = orc_splitmix64_next(&s);
    }
}
