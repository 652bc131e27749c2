/*
 * flock_gpu.h — C ABI of the MI355X (gfx950) implementation of the Flock
 * prover's data-parallel hot path.
 *
 * Every entry point mirrors a reference entry point (cited per declaration,
 * `CF/` = crates/flock-core/src/).  Rust slices become (pointer, length);
 * `&mut [T]` outputs become caller-owned output pointers.  Unless a
 * declaration says "host", every pointer is a DEVICE pointer (HBM-resident
 * buffers, e.g. from hipMalloc or a torch CUDA tensor) and the work is
 * enqueued on `stream` (a hipStream_t; NULL = the default stream) without a
 * host synchronisation.  Nothing retains a pointer after return.
 *
 * Temporary device memory: the entry points that reduce to a few outputs
 * (sumcheck round messages, zerocheck round 1 and round pair, lincheck round
 * message, ring-switch fold_1b_rows) store per-work-group partial sums in
 * scratch taken from the stream-ordered pool of `stream` (hipMallocAsync,
 * released with hipFreeAsync on the same stream); zerocheck round 1 also
 * builds its eq table there, and the opening gathers their index lists.  The
 * proof-of-work calls keep one 8-byte device word and one pinned host word per
 * device.  A failed allocation returns FLK_E_NOMEM.  Calls that return host
 * results (PoW, transcript grinds, read-backs) wait for `stream`; calls that take
 * small HOST arrays (index lists, challenges, constants) stage them in a
 * per-thread pinned upload ring, so the arrays may be reused on return and the
 * call does not wait.
 *
 * Element layout: flk_f128 is 16 bytes, `lo` (x^0..x^63) then `hi`
 * (x^64..x^127), little-endian — identical to the reference F128
 * (CF/field/gf2_128.rs:22-27).  Hashes are 32 raw bytes.
 *
 * Status codes follow the reference's device-offload contract
 * (CF/gpu.rs:13-21,455-466,1219-1221): FLK_E_SHAPE rejects a shape without
 * latching anything; FLK_E_DEVICE is a runtime failure and latches the
 * library disabled until flk_undisable(); FLK_E_DISABLED means the latch,
 * flk_set_enabled(0) or the FLOCK_NO_GPU environment variable is active.
 * There is no CPU fallback inside this library: callers decide.
 */
#ifndef FLOCK_GPU_H
#define FLOCK_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint64_t lo;
    uint64_t hi;
} flk_f128;

enum {
    FLK_OK = 0,
    FLK_E_SHAPE = 1,    /* invalid or unsupported shape; nothing latched */
    FLK_E_DEVICE = 2,   /* HIP runtime failure; library latched disabled */
    FLK_E_DISABLED = 3, /* disabled by latch, API or FLOCK_NO_GPU */
    FLK_E_NOMEM = 4     /* device allocation failed; nothing latched */
};

/* Merkle hash selector, values of the reference HashKind (CF/hash.rs:24-30). */
enum { FLK_HASH_SHA256 = 0, FLK_HASH_BLAKE3 = 1 };

/* ------------------------------------------------------------------------ */
/* Device facade — replaces crate::gpu's metal_available / is_disabled /    */
/* undisable / set_enabled (CF/gpu.rs:67-125).                              */
/* ------------------------------------------------------------------------ */

/* 1 when a HIP device is visible and the library is not disabled. */
int flk_device_available(void);
/* 1 when the failure latch, flk_set_enabled(0) or FLOCK_NO_GPU disables the library. */
int flk_is_disabled(void);
/* Clear the failure latch (CF/gpu.rs undisable). */
void flk_undisable(void);
/* Test hook: after `skip` more entry-point calls, the next n fail as a device
 * runtime failure would (FLK_E_DEVICE, the latch set) without touching the
 * device, so callers can exercise their fallback (CF/gpu.rs:455-516); a skip
 * lands the failure inside a composite call such as flk_blake3_prove.  n = 0
 * cancels. */
void flk_inject_device_failures(unsigned skip, unsigned n);
/* Enable/disable the library (CF/gpu.rs merkle::set_enabled). */
void flk_set_enabled(int enabled);
/* Message for the last non-OK status returned on this thread (host string). */
const char *flk_last_error(void);
/* Library version string (host). */
const char *flk_version(void);
/* Host scalar product in GF(2^128) (F128 Mul, CF/field/gf2_128.rs:140-171), for
 * the host-side bookkeeping of running claims. */
flk_f128 flk_f128_mul(flk_f128 a, flk_f128 b);

/* ------------------------------------------------------------------------ */
/* Device memory for callers without their own HIP binding (a Rust caller   */
/* holding Vec<F128> host buffers).  No reference counterpart: the Metal    */
/* facade wraps host memory zero-copy (CF/gpu.rs:1207-1213), which discrete */
/* HBM cannot; these are the explicit staging calls that replace it.        */
/* ------------------------------------------------------------------------ */

/* hipMalloc of `bytes` into *out (NULL for 0 bytes). */
int flk_alloc(size_t bytes, void **out);
/* Free a pointer from flk_alloc (NULL is a no-op). */
int flk_free(void *p);
/* Stream-ordered pool allocation (hipMallocAsync / hipFreeAsync on `stream`):
 * the memory may be used by work enqueued on `stream` after the allocation and
 * returns to the pool when `stream` reaches the free, without a device-wide
 * synchronisation.  Work on other streams must be ordered against both by the
 * caller.  For the C++ host mirror's temporaries (sumcheck folds, recursive
 * commits); no reference counterpart. */
int flk_alloc_async(size_t bytes, void *stream, void **out);
int flk_free_async(void *p, void *stream);
/* Enqueue a host -> device copy on `stream` (src is HOST memory).  Up to 8 MiB the
 * bytes are staged in the calling thread's pinned upload ring, so src may be
 * reused as soon as the call returns and nothing waits; larger copies wait for
 * the copy. */
int flk_copy_to_device(void *dst, const void *src, size_t bytes, void *stream);
/* Enqueue a device -> host copy on `stream` (dst is HOST memory). */
int flk_copy_to_host(void *dst, const void *src, size_t bytes, void *stream);
/* Copy device -> host and wait for `stream` (dst is HOST memory; small reads go
 * through a per-thread pinned buffer, which halves the latency of a read-back). */
int flk_read_to_host(void *dst, const void *src, size_t bytes, void *stream);
/* Armed read-back of a reduction's result.  flk_read_arm(dst, bytes) marks the
 * next reduction entry point called on this thread whose result is `bytes` long
 * (a round message: flk_zc_fold_round, flk_zc_fold_at_z_round,
 * flk_zc_fold_at_z_bind_round, flk_sc_round_msg, flk_sc_round_msg_and_eval,
 * flk_sc_fold_and_msg, flk_lc_round_msg_msb): the launch that sums its
 * work-group partials also stores the result into the calling thread's pinned
 * mailbox (besides its device output) and sets the mailbox's sequence word, so
 * no publish launch follows it; flk_read_armed(src, bytes, stream) then waits
 * for the word and copies the result to dst.  When no reduction took the arm
 * (another size, an entry point without the fused sum), flk_read_armed reads
 * `src` (the device output) the ordinary way, so the pair is always equivalent
 * to flk_read_to_host(dst, src, bytes).  An arm not followed by flk_read_armed
 * stays on the thread until the next flk_read_arm or flk_read_armed.  No
 * reference counterpart (a latency optimisation of the round read-backs). */
int flk_read_arm(void *dst, size_t bytes);
int flk_read_armed(const void *src, size_t bytes, void *stream);
/* Enqueue a device -> device copy on `stream`. */
int flk_copy_device_to_device(void *dst, const void *src, size_t bytes, void *stream);
/* Wait for all work enqueued on `stream`. */
int flk_synchronize(void *stream);

/* ------------------------------------------------------------------------ */
/* Additive NTT over GF(2^128), LCH novel basis, neighbours-last            */
/* (CF/ntt/additive_ntt_f128.rs).  `dim` is AdditiveNttF128::standard(dim)'s */
/* log domain size; a transform of 2^log_d positions needs log_d <= dim.    */
/* ------------------------------------------------------------------------ */

/* AdditiveNttF128::standard(dim) twiddle table into HOST memory `out`
 * ((2^dim - 1) entries, layer l at offset 2^l - 1; :54-130, :2661-2693). */
int flk_ntt_twiddles(size_t dim, flk_f128 *out);

/* AdditiveNttF128::forward_transform_interleaved_from_layer (:2728-2773):
 * in place on data[pos * num_ntts + lane], n_total = 2^log_d * num_ntts. */
int flk_ntt_forward_interleaved(size_t dim, flk_f128 *data, size_t n_total, size_t num_ntts,
                                size_t start_layer, void *stream);

/* AdditiveNttF128::inverse_transform (:6322-6342), interleaved like the forward
 * transform: in place on data[pos * num_ntts + lane] over all log_d layers,
 * deepest first (v += u; u += v * twiddle).  num_ntts = 1 is inverse_transform. */
int flk_ntt_inverse_interleaved(size_t dim, flk_f128 *data, size_t n_total, size_t num_ntts, void *stream);

/* The pass schedule of the transforms above for n_layers layers over 2^log_lanes
 * interleaved lanes (n_layers = log_d - start_layer; the RS encode runs
 * log_d - log_inv_rate): ks[0..n) layers per pass in forward order, each 1..4,
 * summing to n_layers; n = 0 removes the entry.  Process-wide; shapes without an
 * entry use the library's default (3-layer passes, the remainder as leading
 * 2-layer passes).  The schedule changes only the speed, never a value.  It has
 * no counterpart in the reference, whose CPU schedule is fixed by its top/deep
 * split (:4865-4900); the autotuner of the batched prover
 * (flock_challenge_multi_amd/autotune.py) installs its choices here. */
int flk_ntt_set_schedule(size_t n_layers, size_t log_lanes, const int *ks, size_t n);
/* The schedule in use for (n_layers, log_lanes) into ks[0..*n), *n <= cap. */
int flk_ntt_get_schedule(size_t n_layers, size_t log_lanes, int *ks, size_t cap, size_t *n);

/* AdditiveNttF128::rs_encode_interleaved (:2785-2830): codeword = forward
 * transform of [msg, 0, ...]; n_cw / n_msg is the inverse rate (a power of
 * two >= 2).  msg and codeword must not overlap. */
int flk_ntt_rs_encode_interleaved(size_t dim, const flk_f128 *msg, size_t n_msg, flk_f128 *codeword,
                                  size_t n_cw, size_t num_ntts, void *stream);

/* ------------------------------------------------------------------------ */
/* BLAKE3 Merkle tree (CF/merkle.rs).  Leaf = non-root chaining value of the */
/* leaf bytes; parent = PARENT-flagged merge (:192-203).  Flat 2n-1 layout,  */
/* leaves first, root at index 2n-2 (:15-22).  Leaf sizes: multiples of 16   */
/* from 16 to 1024 bytes (one BLAKE3 chunk).                                 */
/* ------------------------------------------------------------------------ */

/* merkle::hash_leaves (:498): out[i] = leaf CV of data[i*leaf_size ..]. */
int flk_merkle_hash_leaves(const uint8_t *data, size_t leaf_size, size_t n_leaves, uint8_t *out,
                           void *stream);
/* merkle::hash_pairs_level (:683): parents[i] = parent(children[2i], children[2i+1]). */
int flk_merkle_hash_pairs_level(const uint8_t *children, size_t n_children, uint8_t *parents,
                                void *stream);
/* merkle::merkle_tree (:753): tree_out holds 2*num_leaves - 1 hashes. */
int flk_merkle_tree(const uint8_t *data, size_t data_len, size_t num_leaves, int hash_kind,
                    uint8_t *tree_out, void *stream);

/* Streaming Merkle session: crate::gpu::merkle::{begin, Session::{n_leaves,
 * commit_leaves, commit_parent_levels, finish}} (CF/gpu.rs:1228-1490).
 * data (n_leaves * leaf_size bytes) and tree may each be HOST or DEVICE
 * memory; host buffers are mirrored on the device and the nodes the session
 * wrote are copied back by finish.  Parent levels are built while their node
 * count is >= stop_nodes (SIZE_MAX: leaves only; parent levels need a
 * power-of-two leaf count); tree_len (in hashes) must cover every node the
 * session writes.  The caller may write the leaf CVs of ranges it never
 * commits into a HOST tree before commit_parent_levels (the hybrid CPU join,
 * :1161-1170).  All work runs in order on `stream`.  A session is used by one
 * thread at a time. */
typedef struct flk_merkle_session flk_merkle_session;
int flk_merkle_begin(const uint8_t *data, size_t data_len, size_t leaf_size, uint8_t *tree,
                     size_t tree_len, size_t stop_nodes, void *stream, flk_merkle_session **out);
size_t flk_merkle_session_leaves(const flk_merkle_session *s);
/* Hash leaves [leaf_lo, leaf_hi) into tree slots [leaf_lo, leaf_hi); non-blocking. */
int flk_merkle_commit_leaves(flk_merkle_session *s, size_t leaf_lo, size_t leaf_hi);
/* Every parent level with >= stop_nodes nodes; after all leaves are in place. */
int flk_merkle_commit_parent_levels(flk_merkle_session *s);
/* Wait, copy host-tree nodes back, store the session's device time in seconds
 * (gpu_seconds may be NULL) and free the session, on success and on error. */
int flk_merkle_finish(flk_merkle_session *s, double *gpu_seconds);
/* Free a session without finishing it (after waiting for its work). */
void flk_merkle_discard(flk_merkle_session *s);
/* crate::gpu::prewire (CF/gpu.rs:497-515): validated and accepted; this library
 * stages host buffers through HIP copies and registers no host memory. */
int flk_prewire(const void *data, size_t len);

/* ------------------------------------------------------------------------ */
/* PCS commit = RS encode + Merkle over codeword rows.                       */
/* ------------------------------------------------------------------------ */

/* ligerito::ligero_commit (CF/pcs/ligerito.rs:4750-4868) and, with
 * log_num_interleaved = log_batch_size, pcs::commit (CF/pcs/commit.rs:187-227):
 * poly has 2^(log_msg_cols + log_num_interleaved) elements laid out
 * poly[pos * 2^log_num_interleaved + lane]; the codeword has 2^log_inv_rate
 * times as many; tree_out has 2 * 2^(log_msg_cols + log_inv_rate) - 1 hashes and
 * its last entry is the root.  root_out (HOST, may be NULL) receives the root
 * after the stream has been synchronised. */
int flk_commit(const flk_f128 *poly, size_t log_msg_cols, size_t log_num_interleaved,
               size_t log_inv_rate, int hash_kind, flk_f128 *codeword_out, uint8_t *tree_out,
               uint8_t *root_out, void *stream);

/* ------------------------------------------------------------------------ */
/* eq tables and Ligerito sumcheck reductions (CF/pcs/ligerito.rs).          */
/* Message outputs are device arrays of F128: out[0] = u_0, out[1] = u_2.    */
/* ------------------------------------------------------------------------ */

/* lincheck::build_eq_table (CF/lincheck.rs:2080-2121): out[i] = prod_j eq(point[j], bit_j(i)),
 * LSB first.  `point` is a HOST array of d elements; out has 2^d elements. */
int flk_build_eq(const flk_f128 *point, size_t d, flk_f128 *out, void *stream);

/* round_msg_lsb (:4958-5019): u_0 = sum f[2j] b[2j], u_2 = sum (f[2j]+f[2j+1])(b[2j]+b[2j+1]). */
int flk_sc_round_msg(const flk_f128 *f, const flk_f128 *b, size_t n, flk_f128 *out2, void *stream);

/* flk_sc_round_msg without any allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  The work-group partials go to a
 * file-scope device array of 2 x 4096 entries (128 KiB per device, held until
 * the process ends) instead of stream-ordered pool memory; the grid is capped
 * at 4096 work-groups, so every n in [2, 2^40] fits, and larger n is rejected
 * with FLK_E_SHAPE.  All calls on one GPU share that array: calls that overlap
 * on one GPU (other streams, other threads) race and must be serialised by the
 * caller.  Nothing in the array needs to start at zero. */
int flk_sc_round_msg_noalloc(const flk_f128 *f, const flk_f128 *b, size_t n, flk_f128 *out2, void *stream);

/* round_msg_and_eval_lsb (:5020-5075): out3 = {u_0, u_2, sum f.b}. */
int flk_sc_round_msg_and_eval(const flk_f128 *f, const flk_f128 *b, size_t n, flk_f128 *out3,
                              void *stream);

/* fold_and_msg_lsb (:5709-6058) / SumcheckProver::fold (:9012-9070):
 * f_out[j] = f[2j] + r (f[2j] + f[2j+1]), same for b, then the round message of
 * (f_out, b_out) into out2.  n >= 4. */
int flk_sc_fold_and_msg(const flk_f128 *f, const flk_f128 *b, size_t n, flk_f128 r, flk_f128 *f_out,
                        flk_f128 *b_out, flk_f128 *out2, void *stream);

/* SumcheckProver::glue (:9213-9260): b[i] += alpha * b_new[i]. */
int flk_sc_glue(flk_f128 *b, const flk_f128 *b_new, size_t n, flk_f128 alpha, void *stream);

/* partial_eval_lsb (:1162-1175): fold `evals` (n = 2^k' elements) at rs[0..k]
 * (HOST array), LSB first; out has n >> k elements.  `evals` is not modified. */
int flk_partial_eval_lsb(const flk_f128 *evals, size_t n, const flk_f128 *rs, size_t k, flk_f128 *out,
                         void *stream);

/* induce_sumcheck_poly (CF/pcs/ligerito.rs:1436-1548): basis_out (2^log_msg_cols
 * elements) = sum_q alpha_q * X(x_q) with X_j(x) = prod_{k in bits(j)} s_k(x)/s_k(v_k)
 * the LCH basis at x_q = queries[q]; sum_out (1 element) = sum_q alpha_q <rows[q], eq(v)>.
 * rows: device n_queries x 2^n_v matrix (opened codeword rows); queries: device
 * u64[n_queries]; v_challenges (n_v) and alpha (n_alpha, alpha_q = build_eq_table(alpha)[q])
 * are HOST arrays. */
int flk_induce_sumcheck_poly(size_t log_msg_cols, const flk_f128 *rows, size_t n_queries,
                             const flk_f128 *v_challenges, size_t n_v, const uint64_t *queries,
                             const flk_f128 *alpha, size_t n_alpha, flk_f128 *basis_out,
                             flk_f128 *sum_out, void *stream);

/* ------------------------------------------------------------------------ */
/* Lincheck reductions (CF/lincheck.rs).                                     */
/* ------------------------------------------------------------------------ */

/* partial_fold_packed_z / _fast (:2233-2420): z_packed is the lincheck stripe
 * layout of 2^m witness bits (byte (s, i) at z_packed[s * 2^k_log + i], bit r =
 * z[i, 8s + r]); eq_outer has 2^(m - k_log) elements; out[i] = sum over set bits
 * of eq_outer[8s + r], 2^k_log elements.  Needs k_log >= 2, m - k_log >= 3. */
int flk_lc_partial_fold(const uint8_t *z_packed, size_t m, size_t k_log, const flk_f128 *eq_outer,
                        flk_f128 *out, void *stream);
/* Product-sumcheck round over the top variable (:5010-5030): out2 = {q(1), q(inf)} =
 * {sum comb[i+h] z[i+h], sum (comb[i]+comb[i+h])(z[i]+z[i+h])}, h = n/2. */
int flk_lc_round_msg_msb(const flk_f128 *comb, const flk_f128 *z, size_t n, flk_f128 *out2,
                         void *stream);
/* sumcheck_bind_top_in_place (:4968-4990): v[i] += r (v[i+h] + v[i]) for i < h = n/2,
 * in place (the caller then treats v as length h). */
int flk_lc_bind_top(flk_f128 *v, size_t n, flk_f128 r, void *stream);

/* ------------------------------------------------------------------------ */
/* Ring switching (CF/pcs/ring_switch.rs).                                  */
/* ------------------------------------------------------------------------ */

/* fold_1b_rows (:1594-1642): out128[r] = sum_i bit_r(packed_witness[i]) * suffix_tensor[i],
 * r < 128 (polynomial-basis bit r of the element). */
int flk_rs_fold_1b_rows(const flk_f128 *packed_witness, const flk_f128 *suffix_tensor, size_t n,
                        flk_f128 *out128, void *stream);
/* fold_b128_elems (:1952-2087): out[i] = sum_b bit_b(suffix_tensor[i]) * eq_r_dprime[b];
 * eq_r_dprime is a device array of 128 elements. */
int flk_rs_fold_b128_elems(const flk_f128 *suffix_tensor, const flk_f128 *eq_r_dprime, size_t n,
                           flk_f128 *out, void *stream);

struct flk_fs;

/* ring_switch::prove_batched (:3727-3735; transcript order of the per-claim tail
 * :4310-4383) over a dense packed witness of 2^log_len elements (device) and
 * n_claims opening points x_outers (HOST, n_claims rows of log_len + 1
 * coordinates, the reference's x_outer of length m - 6).  For claim k:
 * observe_label("flock-ring-switch-v0"), s_hat_v_k = fold_1b_rows(z, eq(x_outer_k[1..])),
 * observe_f128_slice(s_hat_v_k), r''_k = sample_f128_vec(7),
 * claims_out[k] = <tensor_algebra_transpose(s_hat_v_k), eq(r''_k)>; then
 * gammas_out[k] = sample_f128() for every claim.  s_hat_v_out (HOST) receives
 * n_claims x 128 elements.  b_combined (device, 2^log_len, or NULL) receives
 * sum_k fold_b128_elems(eq(x_outer_k[1..]), gamma_k eq(r''_k)), the sum of the
 * gamma-baked rs_eq_ind that compute_combined_basis_and_target forms
 * (CF/pcs.rs:865-870 and its b_combined sum).  1 <= n_claims <= 64.  The suffix
 * tensors live in stream-ordered scratch for the call (from log_len 20 on only
 * their split factors, 2^12 + 2^(log_len - 12) elements per claim); the call
 * waits for `stream` once, to read s_hat_v.  PaddingSpec only lets the reference
 * skip all-zero chunks; the values equal the dense ones computed here. */
int flk_rs_prove_batched(struct flk_fs *fs, const flk_f128 *packed_witness, size_t log_len,
                         const flk_f128 *x_outers, size_t n_claims, flk_f128 *s_hat_v_out,
                         flk_f128 *claims_out, flk_f128 *gammas_out, flk_f128 *b_combined, void *stream);

/* The out-of-domain samples of a recursive Ligerito level (CF/pcs/ligerito.rs:10569-10590),
 * on the sumcheck pair (f, b) of 2^log_n device elements (1 <= log_n <= 34) and the
 * running target *t_r (HOST, in/out).  Per sample s < n_samples (<= 64):
 * z = sample_f128_vec(log_n); (u_0, u_2, y) = round_msg_and_eval_lsb(f, eq(z))
 * (introduce_new_with_eval, :9144-9154); observe_f128(y), observe_f128(u_0),
 * observe_f128(u_2); beta = sample_f128(); b += beta eq(z); *t_r += beta y (glue,
 * :9213-9256).  ood_values_out[s] = y and msgs_out[2s..2s+2) = (u_0, u_2) are HOST
 * arrays (the proof's ood_values and sumcheck messages).  eq(z) lives in
 * stream-ordered scratch; the call waits for `stream` once per sample, to read
 * (u_0, u_2, y); the last glue stays enqueued. */
int flk_ligerito_ood(struct flk_fs *fs, const flk_f128 *f, flk_f128 *b, size_t log_n, size_t n_samples,
                     flk_f128 *t_r, flk_f128 *ood_values_out, flk_f128 *msgs_out, void *stream);

/* compute_combined_basis_and_target (CF/pcs.rs:778-870) for ring-switched
 * claims (no packed-direct claims): observe_label("flock-pcs-open-batch-v0"),
 * flk_rs_prove_batched, and target_out (HOST) = sum_k gamma_k claims_out[k].
 * b_combined is required.  On a shape error the transcript is untouched. */
int flk_pcs_combined_basis_and_target(struct flk_fs *fs, const flk_f128 *packed_witness, size_t log_len,
                                      const flk_f128 *x_outers, size_t n_claims, flk_f128 *s_hat_v_out,
                                      flk_f128 *claims_out, flk_f128 *gammas_out, flk_f128 *b_combined,
                                      flk_f128 *target_out, void *stream);

/* ------------------------------------------------------------------------ */
/* Zerocheck (CF/zerocheck.rs, CF/zerocheck/multilinear.rs).                */
/* ------------------------------------------------------------------------ */

/* lagrange_weights_naive (multilinear.rs:285-330) into HOST memory: out[i] = L_i(z)
 * for the nodes phi8(0 .. 2^k_skip - 1); k_skip <= 8. */
int flk_zc_lagrange_weights(size_t k_skip, flk_f128 z, flk_f128 *out);
/* fold_at_z (multilinear.rs:575-597): bits holds 2^m witness bits LSB-first, bit
 * x * 2^k_skip + s = a(s, x); weights (device, 2^k_skip) are the Lagrange weights
 * L_s(z); out[x] = sum_s a(s, x) L_s(z), 2^(m - k_skip) elements.  3 <= k_skip <= 6. */
int flk_zc_fold_at_z(const uint8_t *bits, size_t m, size_t k_skip, const flk_f128 *weights,
                     flk_f128 *out, void *stream);
/* Multilinear round message in wire form (zerocheck.rs:1748-1781; the naive
 * round_pair_naive, multilinear.rs:616-640, sends r0 * G(1)):
 * out2 = {G(1), G(inf)} = {sum_x eq_rest[x] a[2x+1] b[2x+1],
 *                          sum_x eq_rest[x] (a[2x]+a[2x+1]) (b[2x]+b[2x+1])}, eq_rest n/2. */
int flk_zc_round_pair(const flk_f128 *a, const flk_f128 *b, size_t n, const flk_f128 *eq_rest,
                      flk_f128 *out2, void *stream);
/* The next round with the previous bind fused in and the eq table split
 * (fold_in_place_pair + round_pair, multilinear.rs:616-640,3194-3213): with
 * rho, T' = fold(T, rho) (T'[j] = T[2j] + rho (T[2j] + T[2j+1]), n inputs, n/2
 * outputs into a_out, b_out); without (rho = NULL) T' = T (n entries).  out2 is
 * the wire message over T' as flk_zc_round_pair computes it with
 * eq_rest[x] = eq_lo[x mod 256] * eq_hi[x >> 8]: eq_lo has 256 entries (the first
 * eight coordinates), eq_hi 2^hi_log, and T' has 2 * 256 * 2^hi_log entries.
 * When hi_log >= 1 and eq_hi_next is not NULL it receives the next round's eq_hi
 * (eq_hi with its lowest coordinate summed out: 2^(hi_log-1) entries). */
int flk_zc_fold_round(const flk_f128 *a, const flk_f128 *b, size_t n, const flk_f128 *rho, const flk_f128 *eq_lo,
                      const flk_f128 *eq_hi, size_t hi_log, flk_f128 *a_out, flk_f128 *b_out, flk_f128 *eq_hi_next,
                      flk_f128 *out2, void *stream);
/* The first multilinear round from the bits (fold_at_z_naive then round_pair,
 * multilinear.rs:575-640) for k_skip = 6: a_out, b_out = the fold at z of the
 * 2^(m-6) rows of 64 bits with the 64 device weights (as flk_zc_fold_at_z), and
 * out2 the message over them as flk_zc_fold_round without rho (hi_log = m - 15;
 * eq_hi_next as there).  With a_out = b_out = NULL the tables are not written:
 * only out2 (and eq_hi_next), as uni_skip_fold_and_round_pair (multilinear.rs:659-683)
 * computes the round over its UniSkipFoldTable (:744-790).  m in 15..40; bit
 * vectors 16-byte aligned. */
int flk_zc_fold_at_z_round(const uint8_t *a_bits, const uint8_t *b_bits, size_t m, const flk_f128 *weights,
                           const flk_f128 *eq_lo, const flk_f128 *eq_hi, flk_f128 *a_out, flk_f128 *b_out,
                           flk_f128 *eq_hi_next, flk_f128 *out2, void *stream);
/* The second multilinear round from the bits: T' = fold(T, rho) with T the fold
 * at z of the rows (fold_at_z_naive then fold_in_place_pair, multilinear.rs:575-597,
 * 3194-3213), 2^(m-7) entries into a_out, b_out, and out2 the message over T' as
 * flk_zc_fold_round computes it (hi_log = m - 16; eq_hi_next as there).  rho is a
 * HOST pointer.  m in 16..40; bit vectors 16-byte aligned. */
int flk_zc_fold_at_z_bind_round(const uint8_t *a_bits, const uint8_t *b_bits, size_t m, const flk_f128 *weights,
                                const flk_f128 *rho, const flk_f128 *eq_lo, const flk_f128 *eq_hi, flk_f128 *a_out,
                                flk_f128 *b_out, flk_f128 *eq_hi_next, flk_f128 *out2, void *stream);

/* Round 1 of the zerocheck (univariate skip over k_skip = 6): round1_naive
 * (univariate_skip.rs:90-152) for the protocol's challenge layout
 * (zerocheck.rs:1680-1692): r[6..9) = phi8(0xF7, 0x53, 0xB5), r[9..13) =
 * x^(2^i) / (1 + x^(2^i)) (univariate_skip_optimized.rs:136-232), r[13..m) =
 * r_outer (HOST array of m - 13 elements).  a_bits, b_bits, c_bits hold 2^m bits
 * each (a_bits, b_bits 8-byte aligned, c_bits 16-byte aligned).  Writes P_AB on Lambda to out_ab[0..64) and P_C to
 * out_c[0..64); out_c must equal out_ab + 64.  13 <= m <= 40. */
int flk_zc_round1(const uint8_t *a_bits, const uint8_t *b_bits, const uint8_t *c_bits, size_t m,
                  const flk_f128 *r_outer, flk_f128 *out_ab, flk_f128 *out_c, void *stream);

/* ------------------------------------------------------------------------ */
/* BLAKE3 R1CS witness (CP/r1cs_hashes/blake3.rs, CP = crates/flock-prover/src/). */
/* ------------------------------------------------------------------------ */

/* One compression input, the reference's Compression tuple (blake3.rs:1860-1861):
 * (cv, m, counter, block_len, flags); 112 bytes. */
typedef struct {
    uint32_t cv[8];
    uint32_t m[16];
    uint64_t counter;
    uint32_t block_len;
    uint32_t flags;
} flk_compression;

/* generate_witness_with_ab_packed_and_lincheck (blake3.rs:4617-4647; with
 * z_lincheck == NULL, generate_witness_with_ab_packed, :2184-2206): for each of
 * the 2^n_blocks_log slots (3 <= n_blocks_log <= 30) the K = 2^14 row values z,
 * a = A z, b = B z of one compression, packed LSB-first, slot j at flk_f128
 * index 128 j of z, a and b (2^n_blocks_log * 128 elements each).  Slots
 * n_blocks.. hold the all-zero compression (blake3.rs:4632-4636).  z_lincheck
 * (2^n_blocks_log * 2048 bytes) receives the lincheck stripe: byte g * 2^14 + p
 * has bit r = z bit p of slot 8 g + r (CP/r1cs_hashes/common.rs:304-320), the
 * layout flk_lc_partial_fold reads.  blocks (device, n_blocks records) and every
 * output must be 16-byte aligned. */
int flk_blake3_witness(const flk_compression *blocks, size_t n_blocks, size_t n_blocks_log, flk_f128 *z,
                       flk_f128 *a, flk_f128 *b, uint8_t *z_lincheck, void *stream);

/* ------------------------------------------------------------------------ */
/* Proof-of-work grinding, BLAKE3 transcripts (CF/challenger.rs).            */
/* ------------------------------------------------------------------------ */

/* pow_scan / blake3_pow_scan (:605-640,1648-1660): *nonce_out = the smallest n in
 * [start, start + len) (the range ends at 2^64 - 1 at most) whose hash
 * blake3(state_digest || u64le(n) || 24 zero bytes) has >= bits leading zero
 * bits (byte 0's most significant bit first), or UINT64_MAX when there is none.
 * state_digest and nonce_out are HOST pointers; synchronises `stream`.  The PoW
 * calls keep an 8-byte device word and a pinned host word per device for the
 * process's life and run one at a time per process. */
int flk_pow_scan(const uint8_t *state_digest, uint64_t start, uint64_t len, uint32_t bits, uint64_t *nonce_out,
                 void *stream);
/* grind_pow's nonce (:384-474): the smallest satisfying nonce from 0, or 0 when
 * bits == 0; 1 <= bits <= 48 otherwise.  The caller absorbs u64le(nonce) into its
 * transcript (observe_bytes), as grind_pow does.  HOST pointers; synchronises. */
int flk_pow_grind(const uint8_t *state_digest, uint32_t bits, uint64_t *nonce_out, void *stream);

/* ------------------------------------------------------------------------ */
/* Fiat-Shamir transcript over BLAKE3, host side (CF/challenger.rs).         */
/* FsChallenger::with_hash(domain, HashKind::Blake3) and its Challenger      */
/* methods (:170-177,226-383,384-474,502-517); every pointer is HOST memory. */
/* ------------------------------------------------------------------------ */
typedef struct flk_fs flk_fs;

/* with_hash(domain, Blake3) (:237-250): absorbs 0x01 || u64le(len) || domain. */
int flk_fs_new(const uint8_t *domain, size_t len, flk_fs **out);
/* Clone (the transcript state is copied). */
int flk_fs_clone(const flk_fs *t, flk_fs **out);
void flk_fs_free(flk_fs *t);
/* observe_label: 0x02 || u64le(len) || label. */
int flk_fs_observe_label(flk_fs *t, const uint8_t *label, size_t len);
/* observe_f128: 0x03 0x01 || lo || hi (little-endian). */
int flk_fs_observe_f128(flk_fs *t, flk_f128 v);
/* observe_f128_slice: 0x03 0x02 || u64le(n) || values. */
int flk_fs_observe_f128_slice(flk_fs *t, const flk_f128 *v, size_t n);
/* observe_bytes: 0x05 || u64le(len) || bytes. */
int flk_fs_observe_bytes(flk_fs *t, const uint8_t *bytes, size_t len);
/* sample_f128 (:355-366): absorb 0x04 0x01, read 16 bytes of the state's XOF,
 * absorb them; the element is (lo, hi) little-endian. */
int flk_fs_sample_f128(flk_fs *t, flk_f128 *out);
/* sample_f128_vec (:368-382): absorb 0x04 0x02 || u64le(n), read 16 n XOF bytes,
 * absorb them. */
int flk_fs_sample_f128_vec(flk_fs *t, size_t n, flk_f128 *out);
/* sample_distinct_queries (CF/pcs/ligerito.rs:9665-9720): sample_f128 until
 * `count` distinct q = lo mod block_len are drawn (first occurrences), written
 * to out sorted ascending; count <= block_len. */
int flk_fs_sample_distinct_queries(flk_fs *t, uint64_t block_len, size_t count, uint64_t *out);
/* state_digest (:309-317): the 32-byte hash of everything absorbed. */
int flk_fs_state_digest(const flk_fs *t, uint8_t out[32]);
/* grind_pow (:384-474): flk_pow_grind on the state digest (device work on
 * `stream`), then observe_bytes(u64le(nonce)); bits == 0 gives nonce 0. */
int flk_fs_grind_pow(flk_fs *t, uint32_t bits, uint64_t *nonce_out, void *stream);
/* verify_pow (:502-523): *ok_out = the PoW check (nonce == 0 when bits == 0);
 * observe_bytes(u64le(nonce)) whatever the outcome.  Host only. */
int flk_fs_verify_pow(flk_fs *t, uint64_t nonce, uint32_t bits, int *ok_out);

/* ------------------------------------------------------------------------ */
/* Ligerito opening glue (CF/merkle.rs, CF/pcs/ligerito.rs).                 */
/* ------------------------------------------------------------------------ */

/* The opened rows: out[i * row_len + j] = codeword[positions[i] * row_len + j]
 * (SoA codeword of n_rows rows of row_len elements, e.g. row_len = num_ntts).
 * positions is a HOST array of n entries < n_rows (staged; the call does not
 * wait). */
int flk_gather_rows(const flk_f128 *codeword, size_t n_rows, size_t row_len, const uint64_t *positions, size_t n,
                    flk_f128 *out, void *stream);
/* merkle_multi_proof (CF/merkle.rs:1020-1060) of a device flat tree (2 num_leaves
 * - 1 hashes): the sibling hashes in emission order into proof_out (device,
 * room for max_hashes), their count in *n_hashes_out (host).  positions is a
 * HOST array (any order, repeats allowed); the sibling indices are staged and
 * the gather enqueued without waiting. */
int flk_merkle_multi_proof(const uint8_t *tree, size_t num_leaves, const uint64_t *positions, size_t n,
                           uint8_t *proof_out, size_t max_hashes, size_t *n_hashes_out, void *stream);

/* ------------------------------------------------------------------------ */
/* The BLAKE3 R1CS and its lincheck circuit (CP/r1cs_hashes/blake3.rs,       */
/* CF/r1cs.rs, CF/lincheck.rs).  A handle holds the block-diagonal instance  */
/* of 2^n_blocks_log compressions (m = 14 + n_blocks_log, k_log = 14,        */
/* k_skip = 6, C_0 = I) and the column-major (A_0, B_0) in HBM (one ~42 MB    */
/* copy per device, shared by every handle).                                 */
/* ------------------------------------------------------------------------ */
typedef struct flk_r1cs flk_r1cs;

/* build_block_r1cs(n_blocks_log) (blake3.rs:1531-1545), 3 <= n_blocks_log <= 26,
 * with its CscCircuit (CF/lincheck.rs:1505-1560) uploaded to the current device. */
int flk_blake3_r1cs_new(size_t n_blocks_log, flk_r1cs **out);
void flk_r1cs_free(flk_r1cs *r);
/* m, k_log, k_skip and the nonzero counts of A_0 and B_0 (any pointer may be NULL). */
int flk_r1cs_shape(const flk_r1cs *r, size_t *m, size_t *k_log, size_t *k_skip, size_t *nnz_a, size_t *nnz_b);
/* BlockR1cs::statement_digest (CF/r1cs.rs:334-385) into HOST out; cached per m. */
int flk_r1cs_statement_digest(const flk_r1cs *r, uint8_t out[32]);
/* The same digest for m = 14 + n_blocks_log without a handle or a device
 * (host only; 17 <= m <= 40). */
int flk_blake3_statement_digest(size_t m, uint8_t out[32]);
/* The lincheck comb vector (CF/lincheck.rs:5590-5660 and the verifier's
 * :5945-5965): comb_out (device, 2^14) = circuit.fold_alpha_batched(alpha,
 * build_quirky_eq_table(z_skip, x_inner_rest, 6)) with beta added at the pinned
 * constant column 512.  x_inner_rest is a HOST array of n_rest = 8 coordinates. */
int flk_lincheck_comb(const flk_r1cs *r, flk_f128 z_skip, const flk_f128 *x_inner_rest, size_t n_rest,
                      flk_f128 alpha, flk_f128 beta, flk_f128 *comb_out, void *stream);

/* ------------------------------------------------------------------------ */
/* Protocol provers sequenced over the device entry points; the transcript   */
/* stays on the host (every output below is HOST memory unless stated).     */
/* ------------------------------------------------------------------------ */

/* zerocheck::prove_packed_padded (CF/zerocheck.rs:384-535; transcript order of
 * verify, :1642-1790) for a(y) b(y) + c(y) = 0 over 2^m points, a_bits, b_bits,
 * c_bits device bit vectors (LSB-first; a_bits, b_bits 8-byte aligned, c_bits
 * 16-byte aligned), 13 <= m <= 40.
 * Outputs: round1_ab[64], round1_c[64] (naive convention), rounds[2 (m - 6)]
 * = (G(1), G(inf)) per multilinear round, finals = {final_a, final_b,
 * final_c}; the claim: *z_out, mlv_out[m - 6] (the rho challenges), r_rest_out[m - 6]. */
int flk_zerocheck_prove(flk_fs *fs, const uint8_t *a_bits, const uint8_t *b_bits, const uint8_t *c_bits, size_t m,
                        flk_f128 *round1_ab, flk_f128 *round1_c, flk_f128 *rounds, flk_f128 *finals,
                        flk_f128 *z_out, flk_f128 *mlv_out, flk_f128 *r_rest_out, void *stream);

/* lincheck::prove_padded (CF/lincheck.rs:5345-5878) with the handle's circuit and
 * constant pin: z_stripe is the device lincheck stripe of the witness
 * (flk_blake3_witness's z_lincheck); x_ab (HOST) = z_skip, x_inner_rest[8],
 * x_outer[m - 14].  Outputs: rounds[16] = (q(1), q(inf)) per round, z_partial[64],
 * the claim *r_inner_skip, r_inner_rest[8] (LSB-first), *w. */
int flk_lincheck_prove(flk_fs *fs, const flk_r1cs *r, const uint8_t *z_stripe, const flk_f128 *x_ab,
                       flk_f128 *rounds, flk_f128 *z_partial, flk_f128 *r_inner_skip, flk_f128 *r_inner_rest,
                       flk_f128 *w, void *stream);

/* Blake3Setup (CP/r1cs_hashes/blake3.rs:4706-4788) for n_blocks compressions:
 * n_blocks_log = ceil(log2(max(n_blocks, 8))), m = 14 + n_blocks_log, PcsParams
 * {m, log_inv_rate of the profile's L0, log_batch_size 6, profile, BLAKE3
 * Merkle}, the Ligerito ladder of (m, profile) (profile 0 fast, 1 slim,
 * 2 secure; ladders exist for m = 22..35, CF/pcs/ligerito.rs:256-345) and the
 * statement digest.  Fast is what the reference's benchmark proves with. */
typedef struct flk_blake3_setup flk_blake3_setup;
int flk_blake3_setup_new(size_t n_blocks, int profile, flk_blake3_setup **out);
void flk_blake3_setup_free(flk_blake3_setup *s);
/* m, n_blocks_log, the device bytes one prove holds at its peak, and the
 * 32-byte statement digest (HOST; any pointer may be NULL). */
int flk_blake3_setup_info(const flk_blake3_setup *s, size_t *m, size_t *n_blocks_log, size_t *device_bytes,
                          uint8_t *digest_out);

/* An upper bound on the bytes flk_blake3_prove writes for setup s (its ladder's
 * opened rows, Merkle multi-proofs, yr, messages and nonces plus the fixed
 * fields): a proof_out buffer of this size never fails for room.  Host-only.
 * No counterpart in the reference, whose to_bytes returns a Vec<u8>
 * (CP/proof_io.rs:160-198). */
int flk_blake3_proof_bound(const flk_blake3_setup *s, size_t *bytes);

/* Device memory of setup s's proves.  Each concurrent prove of s takes a device
 * arena of `capacity` bytes (created at its first use, kept for later proves:
 * no device allocation inside a prove); `peak` is the largest number of bytes
 * any of them has held, `n_arenas` how many exist, `fallbacks` the allocations
 * that did not fit and went to the stream-ordered pool.  Host-only.  The
 * arenas are the device counterpart of the prover's scratch pool, which
 * recycles its large transient buffers across phases and proves
 * (CF/scratch.rs:1-20). */
int flk_blake3_setup_arena_stats(const flk_blake3_setup *s, size_t *capacity, size_t *peak, size_t *n_arenas,
                                 size_t *fallbacks);

/* ---- Batched and served proving (pool.hip) --------------------------------
 * A pool of `workers` host threads (1..16), each with its own HIP stream,
 * proving BLAKE3 R1CS batches of any supported size (m = 22..35) for one
 * Ligerito profile.  Device memory is one arena of `hbm_budget` bytes (0 = 60%
 * of the free device memory): a job starts when its setup's footprint can be
 * carved from it, in submission order.  Setups are made per batch size class on
 * first use (or by flk_prover_pool_prepare).  A job's proof bytes equal
 * flk_blake3_prove's on the same input.  The reference proves one batch per
 * call (CP/prover.rs:373-472) and its harness runs one proof per worker process
 * (benchmark-tools/harness/src/main.rs:360-407); this is the device-side batch
 * and serving form of that loop. */
typedef struct flk_prover_pool flk_prover_pool;
int flk_prover_pool_new(int profile, size_t workers, size_t hbm_budget, flk_prover_pool **out);
/* Running jobs finish; queued jobs are dropped (their waits fail). */
void flk_prover_pool_free(flk_prover_pool *p);
/* Make the setup for batches of n_blocks compressions now (statement digest, R1CS). */
int flk_prover_pool_prepare(flk_prover_pool *p, size_t n_blocks);
/* Queue a proof of n_blocks device compressions (ready on the device, valid until
 * the job is done) under the transcript domain; returns at once with a ticket. */
int flk_prover_pool_submit(flk_prover_pool *p, const flk_compression *blocks, size_t n_blocks, const uint8_t *domain,
                           size_t domain_len, uint64_t *ticket);
/* Wait for a ticket and copy its proof out.  *proof_len is always set; with a
 * buffer too small (or NULL) the call fails with FLK_E_SHAPE and keeps the job.
 * times (optional, 3 doubles): seconds queued, proving, submit to done. */
int flk_prover_pool_wait(flk_prover_pool *p, uint64_t ticket, uint8_t *proof_out, size_t cap, size_t *proof_len,
                         double *times);
int flk_prover_pool_stats(flk_prover_pool *p, size_t *budget, size_t *peak_reserved, size_t *peak_running,
                          size_t *completed);

/* One proof: generate_witness_with_ab_packed_and_lincheck of the device
 * compressions (padding slots are the all-zero compression), pcs::commit,
 * prover::prove_ligerito (CP/prover.rs:373-472: bind_statement, zerocheck,
 * lincheck, ring switch + claim combination, recursive Ligerito open) on a
 * transcript FsChallenger::with_hash(domain, Blake3), serialised as
 * R1csProofBundleLigerito::to_bytes (CP/proof_io.rs:160-198: "FLOCK", version
 * 4, flavor 2, bincode fixint payload) into proof_out (HOST, cap bytes).
 * *proof_len receives the size (also when cap is too small: FLK_E_SHAPE).
 * stage_seconds (HOST, 6 entries, may be NULL) receives the wall time of
 * witness + commit, zerocheck, lincheck, ring switch + combination, Ligerito
 * open and serialisation.  All device work runs on `stream`; concurrent calls
 * on different streams share nothing but the setup (read-only).  Proof-of-work
 * nonces are the smallest valid ones (grind_pow), so the bytes are a function
 * of the inputs. */
int flk_blake3_prove(const flk_blake3_setup *s, const flk_compression *blocks, size_t n_blocks, const uint8_t *domain,
                     size_t domain_len, uint8_t *proof_out, size_t cap, size_t *proof_len, double *stage_seconds,
                     void *stream);

#ifdef __cplusplus
}
#endif
#endif /* FLOCK_GPU_H */
