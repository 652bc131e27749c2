/*
 * CPU oracle for the Flock prover hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, the reference prover's algorithms for the
 * GF(2^128) arithmetic, the additive NTT / Reed-Solomon encode, the BLAKE3
 * Merkle tree, the PCS commit and the sumcheck-family reductions.  It is the
 * checker the GPU path is compared against and the CPU baseline that the
 * benchmark times.  Only the test suite, the smoke check and the benchmark's
 * CPU-baseline leg may load it; the product library never links or calls it.
 *
 * Reference files are cited per function as `CF/...` = crates/flock-core/src/.
 *
 * Element layout (CF/field/gf2_128.rs:22-27): an F128 is 16 bytes, `lo`
 * (coefficients x^0..x^63) then `hi` (x^64..x^127), both little-endian u64.
 */
#ifndef FLOCK_ORACLE_H
#define FLOCK_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { uint64_t lo, hi; } orc_f128;

/* ---- GF(2^128), GHASH polynomial x^128 + x^7 + x^2 + x + 1 ------------- */
orc_f128 orc_f128_mul(orc_f128 a, orc_f128 b);
orc_f128 orc_f128_square(orc_f128 a);
orc_f128 orc_f128_inv(orc_f128 a);
orc_f128 orc_f128_mul_by_x(orc_f128 a);
void orc_clmul64(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi);
void orc_clmul64_bitwise(uint64_t a, uint64_t b, uint64_t *lo, uint64_t *hi);
void orc_f128_mul_unreduced(orc_f128 a, orc_f128 b, uint64_t out[4]);
orc_f128 orc_ghash_reduce(const uint64_t r[4]);
void orc_f128_mul_many(const orc_f128 *a, const orc_f128 *b, orc_f128 *out, size_t n);

/* ---- GF(2^8), AES polynomial, and the phi8 embedding ------------------- */
uint8_t orc_f8_mul(uint8_t a, uint8_t b);
uint8_t orc_f8_inv(uint8_t a);
/* phi8 table derived from the image of x (0x02) by the homomorphism. */
void orc_phi8_table(orc_f128 out[256]);

/* ---- SplitMix64 (reference test RNG, CF/ntt.rs:175-187) ----------------- */
uint64_t orc_splitmix64_next(uint64_t *state);
void orc_fill_f128(uint64_t seed, orc_f128 *out, size_t n);

/* ---- Additive NTT over F128, LCH novel basis (CF/ntt/additive_ntt_f128.rs) */
/* Twiddle table of AdditiveNttF128::standard(dim): (2^dim - 1) entries,
 * layer l at offset 2^l - 1. */
int orc_ntt_twiddles(size_t dim, orc_f128 *out);
int orc_ntt_forward_interleaved(size_t dim, orc_f128 *data, size_t n_total, size_t num_ntts,
                                size_t start_layer);
int orc_ntt_inverse(size_t dim, orc_f128 *data, size_t n);
int orc_ntt_rs_encode_interleaved(size_t dim, const orc_f128 *msg, size_t n_msg,
                                  orc_f128 *codeword, size_t n_cw, size_t num_ntts);
/* Normalised subspace polynomial Ŵ_i evaluated at an arbitrary point (brute
 * force product over the span); used by tests to evaluate the novel basis. */
orc_f128 orc_subspace_poly_hat(size_t i, orc_f128 z);

/* ---- BLAKE3 (specification restatement) -------------------------------- */
void orc_blake3_compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                         uint32_t block_len, uint32_t flags, uint32_t out16[16]);
void orc_blake3_hash(const uint8_t *data, size_t len, uint8_t out[32]);
/* Extended output: out_len bytes of the XOF (the first 32 are the hash). */
void orc_blake3_xof(const uint8_t *data, size_t len, uint8_t *out, size_t out_len);
/* Non-root chaining value of `data` (Hasher::update + finalize_non_root). */
void orc_blake3_leaf_cv(const uint8_t *data, size_t len, uint8_t out[32]);
/* PARENT-flagged non-root merge of two chaining values. */
void orc_blake3_parent_cv(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);
/* Root-flagged parent (used by tests to pin the parent path to BLAKE3 vectors). */
void orc_blake3_parent_root(const uint8_t left[32], const uint8_t right[32], uint8_t out[32]);

/* ---- Merkle tree, flat 2n-1 layout (CF/merkle.rs:15-22, 753-815) -------- */
int orc_merkle_tree(const uint8_t *data, size_t data_len, size_t num_leaves, uint8_t *tree_out);
int orc_merkle_proof(const uint8_t *tree, size_t num_leaves, size_t index, uint8_t *path_out);
int orc_merkle_verify(const uint8_t root[32], const uint8_t leaf_hash[32], size_t index,
                      const uint8_t *path, size_t path_len);

/* ---- PCS commit = RS encode + Merkle (CF/pcs/commit.rs:187-227,1838-1877) */
/* orc_commit on n_threads POSIX threads (same results; commit_mt.c). */
int orc_commit_mt(const orc_f128 *msg, size_t n_msg, size_t num_ntts, size_t log_inv_rate,
                  orc_f128 *codeword_out, uint8_t *tree_out, uint8_t root_out[32], size_t n_threads);
/* Runs fn(ctx, lo, hi) over [0, n) in n_threads contiguous ranges (commit_mt.c). */
int orc_parallel_for(size_t n, size_t n_threads, void (*fn)(void *ctx, size_t lo, size_t hi), void *ctx);
int orc_commit(const orc_f128 *msg, size_t n_msg, size_t num_ntts, size_t log_inv_rate,
               orc_f128 *codeword_out, uint8_t *tree_out, uint8_t root_out[32]);

/* ---- eq tables and multilinear folds ------------------------------------ */
void orc_build_eq(const orc_f128 *point, size_t d, orc_f128 *out);
void orc_partial_eval_lsb(const orc_f128 *evals, size_t n, const orc_f128 *rs, size_t k,
                          orc_f128 *out);

/* ---- Ligerito sumcheck (CF/pcs/ligerito.rs:4958-5019, 5709-6058, 8893-9280) */
void orc_sc_round_msg(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 out[2]);
void orc_sc_fold_and_msg(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 r,
                         orc_f128 *f_out, orc_f128 *b_out, orc_f128 out[2]);
void orc_sc_glue(orc_f128 *b, const orc_f128 *b_new, size_t n, orc_f128 alpha);
void orc_sc_round_msg_and_eval(const orc_f128 *f, const orc_f128 *b, size_t n, orc_f128 out[3]);

/* ---- Ligerito basis induction (CF/pcs/ligerito.rs:1186-1300,1436-1548) -- */
void orc_eval_sk_at_vks(size_t log_n, orc_f128 *out /* log_n + 1 */);
int orc_induce_sumcheck_poly(size_t log_msg_cols, const orc_f128 *rows, size_t n_queries,
                             size_t row_len, const orc_f128 *v_challenges, size_t n_v,
                             const uint64_t *queries, const orc_f128 *alpha, size_t n_alpha,
                             orc_f128 *basis_out, orc_f128 *sum_out);

/* ---- lincheck / ring switch / zerocheck reductions (reductions.c) ------- */
int orc_lc_partial_fold(const uint8_t *z, size_t m, size_t k_log, const orc_f128 *eq_outer, orc_f128 *out);
void orc_lc_bind_top(orc_f128 *v, size_t n, orc_f128 r);
void orc_lc_round_msg_msb(const orc_f128 *comb, const orc_f128 *z, size_t n, orc_f128 out[2]);
void orc_rs_fold_1b_rows(const orc_f128 *z, const orc_f128 *eq, size_t n, orc_f128 out[128]);
void orc_rs_fold_b128_elems(const orc_f128 *suffix, const orc_f128 eq_r[128], size_t n, orc_f128 *out);
int orc_zc_lagrange_weights(size_t k_skip, orc_f128 z, orc_f128 *out);
int orc_zc_fold_at_z(const uint8_t *bits, size_t m, size_t k_skip, const orc_f128 *weights, orc_f128 *out);
void orc_zc_round_pair(const orc_f128 *a, const orc_f128 *b, size_t n, const orc_f128 *eq_rest, orc_f128 out[2]);
/* out[c] = XOR of w[r] over the nonzeros (r, c) of a CSR matrix (transpose times w). */
void orc_sparse_fold(const uint32_t *ptr, const uint32_t *cols, size_t n_rows, const orc_f128 *w, orc_f128 *out,
                     size_t n_cols);

/* ---- GF(2^8) additive NTT and zerocheck round 1 (zerocheck_round1.c) ---- */
int orc_gf8_twiddles(size_t k, uint8_t beta, uint8_t *tw);
int orc_gf8_forward(size_t k, uint8_t beta, uint8_t *v);
int orc_gf8_inverse(size_t k, uint8_t beta, uint8_t *v);
int orc_zc_round1(const uint8_t *a, const uint8_t *b, const uint8_t *c, size_t m, size_t k_skip,
                  const orc_f128 *r, orc_f128 *p_ab, orc_f128 *p_c);
/* orc_zc_round1 for k_skip = 6 on n_threads threads, regrouped sums (zerocheck_round1_mt.c). */
int orc_zc_round1_mt(const uint8_t *a, const uint8_t *b, const uint8_t *c, size_t m, size_t k_skip,
                     const orc_f128 *r, orc_f128 *p_ab, orc_f128 *p_c, size_t n_threads);


/* ---- BLAKE3 R1CS witness generation (witness.c; CP/r1cs_hashes/blake3.rs) */
/* One compression input (blake3.rs:1860-1861): cv, message words, counter,
 * block_len, flags; 112 bytes. */
typedef struct {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} orc_compression;
/* out = {K, const position, G base, G stride, out_hi base, useful bits}. */
void orc_witness_layout(size_t out[6]);
/* (z, a, b) of one block, 256 u64 words each (zeroed first). */
void orc_blake3_block_witness(const orc_compression *blk, uint64_t *z, uint64_t *a, uint64_t *b);
/* 2^n_blocks_log slots (>= 8), padding slots = the all-zero compression; z, a, b
 * hold 2^n_blocks_log * 256 u64 words each; stripe (nullable) 2^n_blocks_log *
 * 2048 bytes. Returns 0, or 1 on a bad shape. */
int orc_blake3_witness(const orc_compression *blocks, size_t n_blocks, size_t n_blocks_log,
                       uint64_t *z, uint64_t *a, uint64_t *b, uint8_t *stripe);

/* Word-form restatement of the same rows (witness_words.c). */
void orc_blake3_block_witness_words(const orc_compression *blk, uint64_t *z, uint64_t *a, uint64_t *b);
/* orc_blake3_witness on n_threads POSIX threads, word form (same outputs). */
int orc_blake3_witness_mt(const orc_compression *blocks, size_t n_blocks, size_t n_blocks_log, uint64_t *z,
                          uint64_t *a, uint64_t *b, uint8_t *stripe, size_t n_threads);

/* ---- Proof-of-work grinding (pow.c; CF/challenger.rs:384-474,549-601) ---- */
/* blake3(state_digest || nonce_le || 24 zero bytes). */
void orc_pow_hash(const uint8_t digest[32], uint64_t nonce, uint8_t out[32]);
/* 1 when h has >= bits leading zero bits (byte 0's most significant bit first). */
int orc_leading_zero_bits_ok(const uint8_t h[32], uint32_t bits);
/* verify_pow's check (bits == 0 accepts only nonce 0). */
int orc_pow_ok(const uint8_t digest[32], uint64_t nonce, uint32_t bits);
/* grind_pow's nonce: the smallest satisfying nonce (0 when bits == 0). */
uint64_t orc_pow_grind(const uint8_t digest[32], uint32_t bits);


/* ---- Ligerito opening glue (open.c; CF/merkle.rs:1020-1191) ------------- */
/* Flat sibling indices of merkle_multi_proof, in emission order; out holds up
 * to n * log2(num_leaves); returns the count. */
size_t orc_merkle_multi_proof_indices(size_t num_leaves, const uint64_t *pos, size_t n, uint64_t *out);
/* verify_merkle_multi_proof over BLAKE3 parents: pos sorted unique. */
int orc_merkle_multi_verify(const uint8_t root[32], size_t num_leaves, const uint64_t *pos, const uint8_t *leaf_hashes,
                            size_t n, const uint8_t *proof, size_t n_proof);
/* out[i] = row pos[i] (row_len elements) of a SoA codeword. */
void orc_gather_rows(const orc_f128 *codeword, size_t row_len, const uint64_t *pos, size_t n, orc_f128 *out);

/* ---- ring switch proving and claim combination (ring_switch.c) --------- */
/* The BLAKE3 Fiat-Shamir transcript as a byte log (CF/challenger.rs:237-383). */
typedef struct orc_fs orc_fs;
orc_fs *orc_fs_new(const uint8_t *domain, size_t len);
void orc_fs_free(orc_fs *t);
int orc_fs_label(orc_fs *t, const uint8_t *label, size_t len);
int orc_fs_bytes(orc_fs *t, const uint8_t *b, size_t len);
int orc_fs_f128(orc_fs *t, orc_f128 v);
int orc_fs_f128_slice(orc_fs *t, const orc_f128 *v, size_t n);
int orc_fs_sample(orc_fs *t, orc_f128 *out);
int orc_fs_sample_vec(orc_fs *t, size_t n, orc_f128 *out);
void orc_fs_digest(const orc_fs *t, uint8_t out[32]);
/* tensor_algebra_transpose: out[b] bit i = bit b of s_hat_v[i]. */
void orc_rs_transpose(const orc_f128 s_hat_v[128], orc_f128 out[128]);
/* ring_switch::prove_batched; b_combined (2^log_len, or NULL) = sum of the
 * gamma-baked rs_eq_ind.  x_outers: n_claims rows of log_len + 1. */
int orc_rs_prove_batched(orc_fs *t, const orc_f128 *z, size_t log_len, const orc_f128 *x_outers, size_t n_claims,
                         orc_f128 *s_hat_v, orc_f128 *claims, orc_f128 *gammas, orc_f128 *b_combined);
int orc_pcs_combined_basis_and_target(orc_fs *t, const orc_f128 *z, size_t log_len, const orc_f128 *x_outers,
                                      size_t n_claims, orc_f128 *s_hat_v, orc_f128 *claims, orc_f128 *gammas,
                                      orc_f128 *b_combined, orc_f128 *target);
/* The verifier's claim check: sum_i build_claim_weights(z_skip, x_outer_0)[i] s_hat_v[i]. */
orc_f128 orc_rs_claim_check(orc_f128 z_skip, orc_f128 x_outer_0, const orc_f128 s_hat_v[128]);
/* The OOD loop of a recursive level (ligerito_ood.c): b and *t_r updated in place;
 * ood_values[n_samples], msgs[2 n_samples] = (u_0, u_2) per sample. */
int orc_ligerito_ood(orc_fs *t, const orc_f128 *f, orc_f128 *b, size_t log_n, size_t n_samples, orc_f128 *t_r,
                     orc_f128 *ood_values, orc_f128 *msgs);

#ifdef __cplusplus
}
#endif
#endif
