/*
 * Oracle (test infrastructure): the lincheck, ring-switch and zerocheck
 * evaluation reductions, restated from the reference's naive forms.
 *
 *   orc_lc_partial_fold     CF/lincheck.rs:2233-2270 (partial_fold_packed_z, stripe layout:
 *                           byte (s, i_inner) at z[s * 2^k_log + i_inner], bit r <-> i_outer = 8s + r)
 *   orc_lc_bind_top         CF/lincheck.rs:4968-4990 (v[i] += r (v[i+half] + v[i]))
 *   orc_lc_round_msg_msb    CF/lincheck.rs:5010-5030 (q(1) = sum hi.zhi, q(inf) = sum (hi+lo)(zhi+zlo))
 *   orc_rs_fold_1b_rows     CF/pcs/ring_switch.rs:1594-1642 (s_hat_v[r] = sum_i bit_r(z_i) eq_i)
 *   orc_rs_fold_b128_elems  CF/pcs/ring_switch.rs:1952-1973 (out_i = sum_b bit_b(suffix_i) eq_r[b])
 *   orc_zc_lagrange_weights CF/zerocheck/multilinear.rs:285-330 (nodes phi8(0..2^k_skip))
 *   orc_zc_fold_at_z        CF/zerocheck/multilinear.rs:575-597
 *   orc_zc_round_pair       CF/zerocheck/multilinear.rs:616-640, returned unscaled as the wire
 *                           pair (G(1), G(inf)) of CF/zerocheck.rs:1748-1781
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static orc_f128 f_add(orc_f128 a, orc_f128 b) {
    orc_f128 o = {a.lo ^ b.lo, a.hi ^ b.hi};
    return o;
}

int orc_lc_partial_fold(const uint8_t *z, size_t m, size_t k_log, const orc_f128 *eq_outer, orc_f128 *out) {
    if (k_log > m || m - k_log < 3) return -1;
    size_t k = (size_t)1 << k_log;
    size_t n_stripes = ((size_t)1 << (m - k_log)) / 8;
    memset(out, 0, k * sizeof(orc_f128));
    for (size_t s = 0; s < n_stripes; s++) {
        const uint8_t *stripe = z + s * k;
        for (size_t i = 0; i < k; i++) {
            uint8_t bits = stripe[i];
            while (bits) {
                int r = __builtin_ctz(bits);
                out[i] = f_add(out[i], eq_outer[8 * s + (size_t)r]);
                bits &= (uint8_t)(bits - 1);
            }
        }
    }
    return 0;
}

void orc_lc_bind_top(orc_f128 *v, size_t n, orc_f128 r) {
    size_t half = n / 2;
    for (size_t i = 0; i < half; i++) v[i] = f_add(v[i], orc_f128_mul(r, f_add(v[i + half], v[i])));
}

void orc_lc_round_msg_msb(const orc_f128 *comb, const orc_f128 *z, size_t n, orc_f128 out[2]) {
    size_t half = n / 2;
    orc_f128 e1 = {0, 0}, einf = {0, 0};
    for (size_t i = 0; i < half; i++) {
        e1 = f_add(e1, orc_f128_mul(comb[i + half], z[i + half]));
        einf = f_add(einf, orc_f128_mul(f_add(comb[i], comb[i + half]), f_add(z[i], z[i + half])));
    }
    out[0] = e1;
    out[1] = einf;
}

void orc_rs_fold_1b_rows(const orc_f128 *z, const orc_f128 *eq, size_t n, orc_f128 out[128]) {
    memset(out, 0, 128 * sizeof(orc_f128));
    for (size_t i = 0; i < n; i++) {
        uint64_t lo = z[i].lo, hi = z[i].hi;
        while (lo) {
            int r = __builtin_ctzll(lo);
            out[r] = f_add(out[r], eq[i]);
            lo &= lo - 1;
        }
        while (hi) {
            int r = __builtin_ctzll(hi);
            out[64 + r] = f_add(out[64 + r], eq[i]);
            hi &= hi - 1;
        }
    }
}

void orc_rs_fold_b128_elems(const orc_f128 *suffix, const orc_f128 eq_r[128], size_t n, orc_f128 *out) {
    for (size_t i = 0; i < n; i++) {
        orc_f128 acc = {0, 0};
        uint64_t lo = suffix[i].lo, hi = suffix[i].hi;
        while (lo) {
            acc = f_add(acc, eq_r[__builtin_ctzll(lo)]);
            lo &= lo - 1;
        }
        while (hi) {
            acc = f_add(acc, eq_r[64 + __builtin_ctzll(hi)]);
            hi &= hi - 1;
        }
        out[i] = acc;
    }
}

int orc_zc_lagrange_weights(size_t k_skip, orc_f128 z, orc_f128 *out) {
    size_t ell = (size_t)1 << k_skip;
    if (ell > 256) return -1;
    orc_f128 phi[256];
    orc_phi8_table(phi);
    for (size_t i = 0; i < ell; i++) {
        orc_f128 num = {1, 0}, den = {1, 0};
        for (size_t j = 0; j < ell; j++) {
            if (j == i) continue;
            num = orc_f128_mul(num, f_add(z, phi[j]));
            den = orc_f128_mul(den, f_add(phi[i], phi[j]));
        }
        out[i] = orc_f128_mul(num, orc_f128_inv(den));
    }
    return 0;
}

int orc_zc_fold_at_z(const uint8_t *bits, size_t m, size_t k_skip, const orc_f128 *weights, orc_f128 *out) {
    if (k_skip > m) return -1;
    size_t ell = (size_t)1 << k_skip;
    size_t n_rest = (size_t)1 << (m - k_skip);
    for (size_t x = 0; x < n_rest; x++) {
        orc_f128 acc = {0, 0};
        for (size_t s = 0; s < ell; s++) {
            size_t idx = x * ell + s;
            if ((bits[idx >> 3] >> (idx & 7)) & 1) acc = f_add(acc, weights[s]);
        }
        out[x] = acc;
    }
    return 0;
}

void orc_zc_round_pair(const orc_f128 *a, const orc_f128 *b, size_t n, const orc_f128 *eq_rest, orc_f128 out[2]) {
    orc_f128 g1 = {0, 0}, ginf = {0, 0};
    for (size_t x = 0; x < n / 2; x++) {
        orc_f128 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1];
        g1 = f_add(g1, orc_f128_mul(eq_rest[x], orc_f128_mul(a1, b1)));
        ginf = f_add(ginf, orc_f128_mul(eq_rest[x], orc_f128_mul(f_add(a0, a1), f_add(b0, b1))));
    }
    out[0] = g1;
    out[1] = ginf;
}

/* out[c] ^= w[r] for every nonzero (r, c) of a CSR matrix with n_rows rows
 * (row r's columns are cols[ptr[r] .. ptr[r+1])): the transpose-times-vector
 * over GF(2^128) that lincheck's fold_alpha_batched applies to A_0 and B_0
 * (CF/lincheck.rs:4683-4720).  out (n_cols entries) is zeroed first. */
void orc_sparse_fold(const uint32_t *ptr, const uint32_t *cols, size_t n_rows, const orc_f128 *w, orc_f128 *out,
                     size_t n_cols)
{
    memset(out, 0, n_cols * sizeof(orc_f128));
    for (size_t r = 0; r < n_rows; ++r)
        for (uint32_t j = ptr[r]; j < ptr[r + 1]; ++j) {
            out[cols[j]].lo ^= w[r].lo;
            out[cols[j]].hi ^= w[r].hi;
        }
}
