/*
 * Ring-switch proving and the PCS claim combination — CPU oracle (TEST
 * INFRASTRUCTURE ONLY; only tests, the smoke check and the benchmark's
 * CPU-baseline leg load it).
 *
 * Restates, over the oracle's own primitives:
 *   - the BLAKE3 Fiat-Shamir transcript as a byte log (CF/challenger.rs:237-250
 *     domain, :327-382 labels, F128 values and slices, squeezes): every squeeze
 *     is the XOF of everything absorbed so far, then re-absorbed; the state
 *     digest is the hash of the log (CF = crates/flock-core/src/);
 *   - ring_switch::prove_batched (CF/pcs/ring_switch.rs:3727-3735, per-claim
 *     tail :4310-4383): label, s_hat_v = fold_1b_rows(z, eq(x_outer[1..])),
 *     observe s_hat_v, r'' = 7 samples, claim = <transpose(s_hat_v), eq(r'')>
 *     (:1646-1680, 1916-1940); then gamma_k for every claim, and
 *     b = sum_k fold_b128_elems(eq(x_outer_k[1..]), gamma_k eq(r''_k));
 *   - compute_combined_basis_and_target (CF/pcs.rs:778-870): the
 *     "flock-pcs-open-batch-v0" label first, target = sum_k gamma_k claim_k;
 *   - the verifier's claim weights (build_claim_weights, ring_switch.rs:138-160)
 *     and claim_check (:1646-1650).
 */
#include <stdlib.h>
#include <string.h>

#include "flock_oracle.h"

struct orc_fs {
    uint8_t *buf;
    size_t len, cap;
};

static int fs_put(orc_fs *t, const void *p, size_t n)
{
    if (t->len + n > t->cap) {
        size_t c = t->cap ? t->cap : 256;
        while (c < t->len + n) c *= 2;
        uint8_t *b = realloc(t->buf, c);
        if (!b) return -1;
        t->buf = b;
        t->cap = c;
    }
    if (n) memcpy(t->buf + t->len, p, n);
    t->len += n;
    return 0;
}

static int fs_u64(orc_fs *t, uint64_t v)
{
    uint8_t b[8];
    for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(v >> (8 * i));
    return fs_put(t, b, 8);
}

static int fs_f128(orc_fs *t, orc_f128 v)
{
    return fs_u64(t, v.lo) || fs_u64(t, v.hi);
}

orc_fs *orc_fs_new(const uint8_t *domain, size_t len)
{
    orc_fs *t = calloc(1, sizeof *t);
    uint8_t tag = 0x01;
    if (!t) return NULL;
    if (fs_put(t, &tag, 1) || fs_u64(t, len) || fs_put(t, domain, len)) {
        orc_fs_free(t);
        return NULL;
    }
    return t;
}

void orc_fs_free(orc_fs *t)
{
    if (t) free(t->buf);
    free(t);
}

int orc_fs_label(orc_fs *t, const uint8_t *label, size_t len)
{
    uint8_t tag = 0x02;
    return fs_put(t, &tag, 1) || fs_u64(t, len) || fs_put(t, label, len);
}

int orc_fs_bytes(orc_fs *t, const uint8_t *b, size_t len)
{
    uint8_t tag = 0x05;
    return fs_put(t, &tag, 1) || fs_u64(t, len) || fs_put(t, b, len);
}

int orc_fs_f128(orc_fs *t, orc_f128 v)
{
    const uint8_t tag[2] = {0x03, 0x01};
    return fs_put(t, tag, 2) || fs_f128(t, v);
}

int orc_fs_f128_slice(orc_fs *t, const orc_f128 *v, size_t n)
{
    const uint8_t tag[2] = {0x03, 0x02};
    if (fs_put(t, tag, 2) || fs_u64(t, n)) return -1;
    for (size_t i = 0; i < n; ++i)
        if (fs_f128(t, v[i])) return -1;
    return 0;
}

static int fs_squeeze(orc_fs *t, size_t n, orc_f128 *out)
{
    uint8_t *b = malloc(16 * n + 1);
    if (!b) return -1;
    orc_blake3_xof(t->buf, t->len, b, 16 * n);
    for (size_t i = 0; i < n; ++i) {
        uint64_t lo = 0, hi = 0;
        for (int j = 0; j < 8; ++j) {
            lo |= (uint64_t)b[16 * i + j] << (8 * j);
            hi |= (uint64_t)b[16 * i + 8 + j] << (8 * j);
        }
        out[i].lo = lo;
        out[i].hi = hi;
    }
    int rc = fs_put(t, b, 16 * n);
    free(b);
    return rc;
}

int orc_fs_sample(orc_fs *t, orc_f128 *out)
{
    const uint8_t tag[2] = {0x04, 0x01};
    return fs_put(t, tag, 2) || fs_squeeze(t, 1, out);
}

int orc_fs_sample_vec(orc_fs *t, size_t n, orc_f128 *out)
{
    const uint8_t tag[2] = {0x04, 0x02};
    return fs_put(t, tag, 2) || fs_u64(t, n) || fs_squeeze(t, n, out);
}

void orc_fs_digest(const orc_fs *t, uint8_t out[32])
{
    orc_blake3_hash(t->buf, t->len, out);
}

void orc_rs_transpose(const orc_f128 s_hat_v[128], orc_f128 out[128])
{
    memset(out, 0, 128 * sizeof *out);
    for (unsigned i = 0; i < 128; ++i)
        for (unsigned b = 0; b < 128; ++b) {
            uint64_t w = b < 64 ? s_hat_v[i].lo : s_hat_v[i].hi;
            if ((w >> (b & 63)) & 1) {
                if (i < 64) out[b].lo |= 1ull << i;
                else out[b].hi |= 1ull << (i - 64);
            }
        }
}

static orc_f128 inner128(const orc_f128 *a, const orc_f128 *b)
{
    orc_f128 acc = {0, 0};
    for (unsigned i = 0; i < 128; ++i) {
        orc_f128 p = orc_f128_mul(a[i], b[i]);
        acc.lo ^= p.lo;
        acc.hi ^= p.hi;
    }
    return acc;
}

static const char RS_LABEL[] = "flock-ring-switch-v0";
static const char PCS_LABEL[] = "flock-pcs-open-batch-v0";

int orc_rs_prove_batched(orc_fs *t, const orc_f128 *z, size_t log_len, const orc_f128 *x_outers, size_t n_claims,
                         orc_f128 *s_hat_v, orc_f128 *claims, orc_f128 *gammas, orc_f128 *b_combined)
{
    const size_t n = (size_t)1 << log_len, d = log_len + 1;
    if (n_claims == 0) return -1;
    orc_f128 *sfx = malloc(n * sizeof *sfx), *eq_r = malloc(n_claims * 128 * sizeof *eq_r);
    orc_f128 *tmp = malloc(n * sizeof *tmp);
    int rc = (!sfx || !eq_r || !tmp) ? -1 : 0;
    for (size_t k = 0; k < n_claims && !rc; ++k) {
        orc_build_eq(x_outers + k * d + 1, log_len, sfx);
        orc_rs_fold_1b_rows(z, sfx, n, s_hat_v + 128 * k);
    }
    for (size_t k = 0; k < n_claims && !rc; ++k) {
        orc_f128 r[7], u[128];
        rc = orc_fs_label(t, (const uint8_t *)RS_LABEL, sizeof RS_LABEL - 1) ||
             orc_fs_f128_slice(t, s_hat_v + 128 * k, 128) || orc_fs_sample_vec(t, 7, r);
        if (rc) break;
        orc_build_eq(r, 7, eq_r + 128 * k);
        orc_rs_transpose(s_hat_v + 128 * k, u);
        claims[k] = inner128(u, eq_r + 128 * k);
    }
    for (size_t k = 0; k < n_claims && !rc; ++k) rc = orc_fs_sample(t, gammas + k);
    if (!rc && b_combined) {
        memset(b_combined, 0, n * sizeof *b_combined);
        for (size_t k = 0; k < n_claims; ++k) {
            orc_f128 eg[128];
            for (unsigned b = 0; b < 128; ++b) eg[b] = orc_f128_mul(gammas[k], eq_r[128 * k + b]);
            orc_build_eq(x_outers + k * d + 1, log_len, sfx);
            orc_rs_fold_b128_elems(sfx, eg, n, tmp);
            for (size_t i = 0; i < n; ++i) {
                b_combined[i].lo ^= tmp[i].lo;
                b_combined[i].hi ^= tmp[i].hi;
            }
        }
    }
    free(sfx);
    free(eq_r);
    free(tmp);
    return rc;
}

int orc_pcs_combined_basis_and_target(orc_fs *t, const orc_f128 *z, size_t log_len, const orc_f128 *x_outers,
                                      size_t n_claims, orc_f128 *s_hat_v, orc_f128 *claims, orc_f128 *gammas,
                                      orc_f128 *b_combined, orc_f128 *target)
{
    if (n_claims == 0) return -1;
    if (orc_fs_label(t, (const uint8_t *)PCS_LABEL, sizeof PCS_LABEL - 1)) return -1;
    if (orc_rs_prove_batched(t, z, log_len, x_outers, n_claims, s_hat_v, claims, gammas, b_combined)) return -1;
    orc_f128 acc = {0, 0};
    for (size_t k = 0; k < n_claims; ++k) {
        orc_f128 p = orc_f128_mul(gammas[k], claims[k]);
        acc.lo ^= p.lo;
        acc.hi ^= p.hi;
    }
    *target = acc;
    return 0;
}

orc_f128 orc_rs_claim_check(orc_f128 z_skip, orc_f128 x_outer_0, const orc_f128 s_hat_v[128])
{
    orc_f128 lambda[64], w[128];
    orc_zc_lagrange_weights(6, z_skip, lambda);
    orc_f128 eq_lo = {x_outer_0.lo ^ 1, x_outer_0.hi};
    for (unsigned i = 0; i < 128; ++i) w[i] = orc_f128_mul(lambda[i & 63], (i >> 6) & 1 ? x_outer_0 : eq_lo);
    return inner128(w, s_hat_v);
}
