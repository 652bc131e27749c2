/*
 * Oracle (test infrastructure): additive NTT over GF(2^128) in the LCH novel
 * polynomial basis, neighbours-last ordering, interleaved (SoA) layout.
 *
 *   subspace evals        CF/ntt/additive_ntt_f128.rs:54-84 (generate_evals_from_subspace)
 *   twiddle(l, b)         CF/ntt/additive_ntt_f128.rs:2684-2693 (span of the eval row)
 *   forward interleaved   CF/ntt/additive_ntt_f128.rs:4766-4795 (scalar, from start_layer)
 *   inverse               CF/ntt/additive_ntt_f128.rs:6322-6342
 *   rs_encode_interleaved CF/ntt/additive_ntt_f128.rs:2785-2830, defined as zero-padding the
 *                         message and running the full forward transform
 *                         (CF/pcs/commit.rs:1838-1877 commit_matches_full_ntt_oracle)
 *
 * The standard basis is beta_i = x^i (AdditiveNttF128::standard, :2661-2670).
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

static orc_f128 f_add(orc_f128 a, orc_f128 b) {
    orc_f128 o = {a.lo ^ b.lo, a.hi ^ b.hi};
    return o;
}

/* evals[i][k] = Ŵ_i(beta_{i+k}), k = 0..dim-1-i, stored row-major with row i
 * at offset i*dim. */
static orc_f128 *subspace_evals(size_t dim) {
    orc_f128 *ev = (orc_f128 *)calloc(dim * dim + 1, sizeof(orc_f128));
    if (!ev) return NULL;
    for (size_t k = 0; k < dim; k++) {
        ev[k].lo = k < 64 ? (1ull << k) : 0;
        ev[k].hi = k >= 64 ? (1ull << (k - 64)) : 0;
    }
    for (size_t i = 1; i < dim; i++) {
        const orc_f128 *prev = ev + (i - 1) * dim;
        orc_f128 *row = ev + i * dim;
        size_t prev_len = dim - (i - 1);
        for (size_t k = 1; k < prev_len; k++)
            row[k - 1] = orc_f128_mul(prev[k], f_add(prev[k], prev[0]));
    }
    for (size_t i = 0; i < dim; i++) {
        orc_f128 *row = ev + i * dim;
        orc_f128 inv = orc_f128_inv(row[0]);
        for (size_t k = 0; k < dim - i; k++) row[k] = orc_f128_mul(row[k], inv);
    }
    return ev;
}

int orc_ntt_twiddles(size_t dim, orc_f128 *out) {
    if (dim == 0) return 0;
    if (dim > 64) return -1;
    orc_f128 *ev = subspace_evals(dim);
    if (!ev) return -2;
    for (size_t layer = 0; layer < dim; layer++) {
        const orc_f128 *row = ev + (dim - layer - 1) * dim; /* length layer+1 */
        size_t base = ((size_t)1 << layer) - 1;
        for (size_t b = 0; b < ((size_t)1 << layer); b++) {
            orc_f128 acc = {0, 0};
            for (size_t j = 0; j < layer; j++)
                if ((b >> j) & 1) acc = f_add(acc, row[1 + j]);
            out[base + b] = acc;
        }
    }
    free(ev);
    return 0;
}

static int log2_exact(size_t n) {
    if (n == 0 || (n & (n - 1))) return -1;
    int l = 0;
    while (((size_t)1 << l) < n) l++;
    return l;
}

int orc_ntt_forward_interleaved(size_t dim, orc_f128 *data, size_t n_total, size_t num_ntts,
                                size_t start_layer) {
    if (num_ntts == 0 || (num_ntts & (num_ntts - 1)) || n_total % num_ntts) return -1;
    int log_d = log2_exact(n_total / num_ntts);
    if (log_d < 0 || (size_t)log_d > dim || start_layer > (size_t)log_d) return -1;
    if (dim == 0) return 0;
    orc_f128 *tw = (orc_f128 *)malloc((((size_t)1 << dim) - 1) * sizeof(orc_f128));
    if (!tw) return -2;
    orc_ntt_twiddles(dim, tw);
    for (size_t layer = start_layer; layer < (size_t)log_d; layer++) {
        size_t num_blocks = (size_t)1 << layer;
        size_t block_size = (size_t)1 << (log_d - layer);
        size_t half = block_size >> 1;
        for (size_t block = 0; block < num_blocks; block++) {
            orc_f128 t = tw[num_blocks - 1 + block];
            size_t block_start = block * block_size * num_ntts;
            for (size_t row = 0; row < half; row++) {
                size_t off_top = block_start + row * num_ntts;
                size_t off_bot = off_top + half * num_ntts;
                for (size_t lane = 0; lane < num_ntts; lane++) {
                    orc_f128 v = data[off_bot + lane];
                    orc_f128 u = f_add(data[off_top + lane], orc_f128_mul(v, t));
                    data[off_top + lane] = u;
                    data[off_bot + lane] = f_add(v, u);
                }
            }
        }
    }
    free(tw);
    return 0;
}

int orc_ntt_inverse(size_t dim, orc_f128 *data, size_t n) {
    int log_d = log2_exact(n);
    if (log_d < 0 || (size_t)log_d > dim) return -1;
    if (dim == 0) return 0;
    orc_f128 *tw = (orc_f128 *)malloc((((size_t)1 << dim) - 1) * sizeof(orc_f128));
    if (!tw) return -2;
    orc_ntt_twiddles(dim, tw);
    for (int layer = log_d - 1; layer >= 0; layer--) {
        size_t num_blocks = (size_t)1 << layer;
        size_t half = (size_t)1 << (log_d - layer - 1);
        for (size_t block = 0; block < num_blocks; block++) {
            orc_f128 t = tw[num_blocks - 1 + block];
            size_t start = block << (log_d - layer);
            for (size_t i0 = start; i0 < start + half; i0++) {
                size_t i1 = i0 | half;
                orc_f128 u = data[i0];
                orc_f128 nv = f_add(data[i1], u);
                data[i1] = nv;
                data[i0] = f_add(u, orc_f128_mul(nv, t));
            }
        }
    }
    free(tw);
    return 0;
}

int orc_ntt_rs_encode_interleaved(size_t dim, const orc_f128 *msg, size_t n_msg,
                                  orc_f128 *codeword, size_t n_cw, size_t num_ntts) {
    if (n_msg == 0 || num_ntts == 0 || n_msg % num_ntts || n_cw % n_msg) return -1;
    size_t inv_rate = n_cw / n_msg;
    if (inv_rate < 2 || (inv_rate & (inv_rate - 1))) return -1;
    memset(codeword, 0, n_cw * sizeof(orc_f128));
    memcpy(codeword, msg, n_msg * sizeof(orc_f128));
    return orc_ntt_forward_interleaved(dim, codeword, n_cw, num_ntts, 0);
}

orc_f128 orc_subspace_poly_hat(size_t i, orc_f128 z) {
    /* W_i(z) = prod_{v in span(x^0..x^{i-1})} (z + v), normalised by W_i(x^i). */
    orc_f128 num = {1, 0}, den = {1, 0};
    orc_f128 bi = {i < 64 ? (1ull << i) : 0, i >= 64 ? (1ull << (i - 64)) : 0};
    for (size_t m = 0; m < ((size_t)1 << i); m++) {
        orc_f128 v = {(uint64_t)m, 0};
        num = orc_f128_mul(num, f_add(z, v));
        den = orc_f128_mul(den, f_add(bi, v));
    }
    return orc_f128_mul(num, orc_f128_inv(den));
}
