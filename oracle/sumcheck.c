/*
 * Oracle (test infrastructure): eq tables, LSB-first multilinear folds and the
 * Ligerito sumcheck reductions.
 *
 *   orc_build_eq             CF/lincheck.rs:2080-2121 (build_eq_table: LSB-first doubling,
 *                            out[i] = v + v*r_j, out[i+len] = v*r_j)
 *   orc_partial_eval_lsb     CF/pcs/ligerito.rs:1162-1175
 *   orc_sc_round_msg         CF/pcs/ligerito.rs:4958-5019 (u0 = sum f0 b0,
 *                            u2 = sum (f0+f1)(b0+b1))
 *   orc_sc_round_msg_and_eval CF/pcs/ligerito.rs:5020-5075 (adds y = sum f.b)
 *   orc_sc_fold_and_msg      CF/pcs/ligerito.rs:5709-6058 (fold f and b at r, then the
 *                            round message of the folded pair)
 *   orc_sc_glue              CF/pcs/ligerito.rs:9213-9260 (b += alpha * b_new)
 *   orc_eval_sk_at_vks       CF/pcs/ligerito.rs:1186-1205
 *   orc_induce_sumcheck_poly CF/pcs/ligerito.rs:1436-1548 with
 *                            evaluate_scaled_basis_inplace :1237-1300
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static orc_f128 f_add(orc_f128 a, orc_f128 b) {
    orc_f128 o = {a.lo ^ b.lo, a.hi ^ b.hi};
    return o;
}

void orc_build_eq(const orc_f128 *point, size_t d, orc_f128 *out) {
    out[0].lo = 1;
    out[0].hi = 0;
    for (size_t j = 0; j < d; j++) {
        size_t len = (size_t)1 << j;
        for (size_t i = 0; i < len; i++) {
            orc_f128 v = out[i];
            orc_f128 hi = orc_f128_mul(v, point[j]);
            out[i + len] = hi;
            out[i] = f_add(v, hi);
        }
    }
}

void orc_partial_eval_lsb(const orc_f128 *evals, size_t n, const orc_f128 *rs, size_t k,
                          orc_f128 *out) {
    orc_f128 *cur = (orc_f128 *)malloc(n * sizeof(orc_f128));
    memcpy(cur, evals, n * sizeof(orc_f128));
    size_t len = n;
    for (size_t t = 0; t < k; t++) {
        size_t half = len / 2;
        for (size_t i = 0; i < half; i++) {
            orc_f128 e0 = cur[2 * i], e1 = cur[2 * i + 1];
            cur[i] = f_add(e0, orc_f128_mul(rs[t], f_add(e0, e1)));
        }
        len = half;
    }
    memcpy(out, cur, len * sizeof(orc_f128));
    free(cur);
}

void orc_sc_round_msg(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 out[2]) {
    orc_f128 u0 = {0, 0}, u2 = {0, 0};
    for (size_t j = 0; j < n / 2; j++) {
        orc_f128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        u0 = f_add(u0, orc_f128_mul(f0, b0));
        u2 = f_add(u2, orc_f128_mul(f_add(f0, f1), f_add(b0, b1)));
    }
    out[0] = u0;
    out[1] = u2;
}

void orc_sc_round_msg_and_eval(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 out[3]) {
    orc_f128 u0 = {0, 0}, u2 = {0, 0}, y = {0, 0};
    for (size_t j = 0; j < n / 2; j++) {
        orc_f128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        orc_f128 e0 = orc_f128_mul(f0, b0);
        u0 = f_add(u0, e0);
        u2 = f_add(u2, orc_f128_mul(f_add(f0, f1), f_add(b0, b1)));
        y = f_add(y, f_add(e0, orc_f128_mul(f1, b1)));
    }
    out[0] = u0;
    out[1] = u2;
    out[2] = y;
}

void orc_sc_fold_and_msg(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 r,
                         orc_f128 *f_out, orc_f128 *b_out, orc_f128 out[2]) {
    size_t half = n / 2;
    for (size_t j = 0; j < half; j++) {
        orc_f128 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
        f_out[j] = f_add(f0, orc_f128_mul(r, f_add(f0, f1)));
        b_out[j] = f_add(b0, orc_f128_mul(r, f_add(b0, b1)));
    }
    orc_sc_round_msg(f_out, b_out, half, out);
}

void orc_sc_glue(orc_f128 *b, const orc_f128 *b_new, size_t n, orc_f128 alpha) {
    for (size_t i = 0; i < n; i++) b[i] = f_add(b[i], orc_f128_mul(alpha, b_new[i]));
}

static orc_f128 next_s(orc_f128 s, orc_f128 s_at_root) {
    return f_add(orc_f128_mul(s, s), orc_f128_mul(s_at_root, s));
}

void orc_eval_sk_at_vks(size_t log_n, orc_f128 *sks) {
    for (size_t i = 0; i <= log_n; i++) sks[i].lo = sks[i].hi = 0;
    sks[0].lo = 1;
    if (log_n == 0) return;
    orc_f128 *layer = (orc_f128 *)malloc(log_n * sizeof(orc_f128));
    for (size_t i = 1; i <= log_n; i++) {
        layer[i - 1].lo = 1ull << i;
        layer[i - 1].hi = 0;
    }
    size_t cur_len = log_n;
    for (size_t i = 0; i < log_n; i++) {
        for (size_t j = 0; j < cur_len; j++) {
            orc_f128 v = next_s(layer[j], sks[i]);
            if (j == 0)
                sks[i + 1] = v;
            else
                layer[j - 1] = v;
        }
        cur_len--;
    }
    free(layer);
}

int orc_induce_sumcheck_poly(size_t log_msg_cols, const orc_f128 *rows, size_t n_queries,
                             size_t row_len, const orc_f128 *v_challenges, size_t n_v,
                             const uint64_t *queries, const orc_f128 *alpha, size_t n_alpha,
                             orc_f128 *basis_out, orc_f128 *sum_out) {
    size_t n = (size_t)1 << log_msg_cols;
    if (row_len != ((size_t)1 << n_v)) return -1;
    if (n_queries > ((size_t)1 << n_alpha)) return -1;
    orc_f128 *eq = (orc_f128 *)malloc(row_len * sizeof(orc_f128));
    orc_build_eq(v_challenges, n_v, eq);
    orc_f128 *apows = (orc_f128 *)malloc(((size_t)1 << n_alpha) * sizeof(orc_f128));
    orc_build_eq(alpha, n_alpha, apows);
    orc_f128 *sks_vks = (orc_f128 *)malloc((log_msg_cols + 1) * sizeof(orc_f128));
    orc_eval_sk_at_vks(log_msg_cols, sks_vks);
    orc_f128 *inv_sks = (orc_f128 *)malloc((log_msg_cols + 1) * sizeof(orc_f128));
    for (size_t i = 0; i <= log_msg_cols; i++) {
        if (sks_vks[i].lo == 0 && sks_vks[i].hi == 0)
            inv_sks[i] = sks_vks[i];
        else
            inv_sks[i] = orc_f128_inv(sks_vks[i]);
    }
    orc_f128 *local = (orc_f128 *)malloc(n * sizeof(orc_f128));
    orc_f128 *sks_at_x = (orc_f128 *)malloc((log_msg_cols + 1) * sizeof(orc_f128));
    memset(basis_out, 0, n * sizeof(orc_f128));
    orc_f128 sum = {0, 0};
    for (size_t qi = 0; qi < n_queries; qi++) {
        const orc_f128 *row = rows + qi * row_len;
        orc_f128 dot = {0, 0};
        for (size_t k = 0; k < row_len; k++) dot = f_add(dot, orc_f128_mul(row[k], eq[k]));
        orc_f128 ap = apows[qi];
        sum = f_add(sum, orc_f128_mul(dot, ap));
        /* evaluate_scaled_basis_inplace(x = q, alpha = ap) */
        orc_f128 x = {queries[qi], 0};
        if (log_msg_cols > 0) {
            sks_at_x[0] = x;
            for (size_t i = 1; i < log_msg_cols; i++) sks_at_x[i] = next_s(sks_at_x[i - 1], sks_vks[i - 1]);
            for (size_t i = 0; i < log_msg_cols; i++) sks_at_x[i] = orc_f128_mul(sks_at_x[i], inv_sks[i]);
        }
        local[0] = ap;
        for (size_t k = 0; k < log_msg_cols; k++) {
            size_t len = (size_t)1 << k;
            for (size_t i = 0; i < len; i++) local[i + len] = orc_f128_mul(sks_at_x[k], local[i]);
        }
        for (size_t i = 0; i < n; i++) basis_out[i] = f_add(basis_out[i], local[i]);
    }
    *sum_out = sum;
    free(eq);
    free(apows);
    free(sks_vks);
    free(inv_sks);
    free(local);
    free(sks_at_x);
    return 0;
}
