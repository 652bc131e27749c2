/*
 * Oracle (test infrastructure): the GF(2^8) additive NTT and the zerocheck
 * round-1 (univariate skip) message.
 *
 *   orc_gf8_twiddles   CF/ntt.rs:33-69 (compute_twiddles(k, beta), heap layout tw[idx-1])
 *   orc_gf8_forward    CF/ntt.rs:71-93,150-158 (fft_rec)
 *   orc_gf8_inverse    CF/ntt.rs:95-117,160-168 (ifft_rec)
 *   orc_zc_round1      CF/zerocheck/univariate_skip.rs:90-152 (round1_naive):
 *     for each x_rest: the 2^k_skip bits a(s, x) on S = {0..2^k_skip-1} are
 *     extended to Lambda = {2^k_skip ..} by inverse NTT on S then forward NTT on
 *     Lambda; P_AB[l] += eq(r_rest, x) phi8(a_l b_l), P_C[l] += eq(r_rest, x) phi8(c_l).
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static uint8_t next_s8(uint8_t s, uint8_t root) { return orc_f8_mul(s, s) ^ orc_f8_mul(root, s); }

int orc_gf8_twiddles(size_t k, uint8_t beta, uint8_t *tw) {
    if (k == 0) return 0;
    if (k > 8) return -1;
    size_t len = (size_t)1 << (k - 1);
    uint8_t layer[128];
    for (size_t i = 0; i < len; i++) layer[i] = beta ^ (uint8_t)(2 * i);
    uint8_t s_at_root = 1;
    size_t write_at = len;
    for (size_t i = 0; i < len; i++) tw[write_at - 1 + i] = layer[i];
    for (size_t it = 1; it < k; it++) {
        write_at >>= 1;
        uint8_t next_root = next_s8(layer[1] ^ layer[0], s_at_root);
        size_t new_len = write_at;
        for (size_t i = 0; i < new_len; i++) layer[i] = next_s8(layer[2 * i], s_at_root);
        len = new_len;
        s_at_root = next_root;
        uint8_t s_inv = orc_f8_inv(s_at_root);
        for (size_t j = 0; j < len; j++) tw[write_at - 1 + j] = orc_f8_mul(s_inv, layer[j]);
    }
    return 0;
}

static void fft_rec(uint8_t *v, size_t n, const uint8_t *tw, size_t idx) {
    if (n == 1) return;
    size_t half = n >> 1;
    uint8_t lam = tw[idx - 1];
    for (size_t i = 0; i < half; i++) {
        uint8_t w = v[half + i];
        v[i] ^= orc_f8_mul(lam, w);
        v[half + i] = w ^ v[i];
    }
    fft_rec(v, half, tw, 2 * idx);
    fft_rec(v + half, half, tw, 2 * idx + 1);
}

static void ifft_rec(uint8_t *v, size_t n, const uint8_t *tw, size_t idx) {
    if (n == 1) return;
    size_t half = n >> 1;
    ifft_rec(v, half, tw, 2 * idx);
    ifft_rec(v + half, half, tw, 2 * idx + 1);
    uint8_t lam = tw[idx - 1];
    for (size_t i = 0; i < half; i++) {
        v[half + i] ^= v[i];
        v[i] ^= orc_f8_mul(lam, v[half + i]);
    }
}

int orc_gf8_forward(size_t k, uint8_t beta, uint8_t *v) {
    uint8_t tw[256];
    if (orc_gf8_twiddles(k, beta, tw)) return -1;
    if (k) fft_rec(v, (size_t)1 << k, tw, 1);
    return 0;
}

int orc_gf8_inverse(size_t k, uint8_t beta, uint8_t *v) {
    uint8_t tw[256];
    if (orc_gf8_twiddles(k, beta, tw)) return -1;
    if (k) ifft_rec(v, (size_t)1 << k, tw, 1);
    return 0;
}

int orc_zc_round1(const uint8_t *a, const uint8_t *b, const uint8_t *c, size_t m, size_t k_skip,
                  const orc_f128 *r, orc_f128 *p_ab, orc_f128 *p_c) {
    if (k_skip > m || k_skip > 7) return -1;
    size_t ell = (size_t)1 << k_skip;
    size_t nx = (size_t)1 << (m - k_skip);
    uint8_t twS[256], twL[256];
    orc_gf8_twiddles(k_skip, 0, twS);
    orc_gf8_twiddles(k_skip, (uint8_t)ell, twL);
    orc_f128 *eq = (orc_f128 *)malloc(nx * sizeof(orc_f128));
    orc_build_eq(r + k_skip, m - k_skip, eq);
    orc_f128 phi[256];
    orc_phi8_table(phi);
    memset(p_ab, 0, ell * sizeof(orc_f128));
    memset(p_c, 0, ell * sizeof(orc_f128));
    uint8_t ac[128], bc[128], cc[128];
    for (size_t x = 0; x < nx; x++) {
        for (size_t s = 0; s < ell; s++) {
            size_t idx = x * ell + s;
            ac[s] = (a[idx >> 3] >> (idx & 7)) & 1;
            bc[s] = (b[idx >> 3] >> (idx & 7)) & 1;
            cc[s] = (c[idx >> 3] >> (idx & 7)) & 1;
        }
        if (k_skip) {
            ifft_rec(ac, ell, twS, 1);
            fft_rec(ac, ell, twL, 1);
            ifft_rec(bc, ell, twS, 1);
            fft_rec(bc, ell, twL, 1);
            ifft_rec(cc, ell, twS, 1);
            fft_rec(cc, ell, twL, 1);
        }
        for (size_t i = 0; i < ell; i++) {
            orc_f128 t1 = orc_f128_mul(eq[x], phi[orc_f8_mul(ac[i], bc[i])]);
            orc_f128 t2 = orc_f128_mul(eq[x], phi[cc[i]]);
            p_ab[i].lo ^= t1.lo;
            p_ab[i].hi ^= t1.hi;
            p_c[i].lo ^= t2.lo;
            p_c[i].hi ^= t2.hi;
        }
    }
    free(eq);
    return 0;
}
