// Host-side helpers of the device library for the protocol code compiled
// without the HIP headers (the .cpp files): status reporting, host GF(2^128)
// arithmetic and the zerocheck constants (CF = crates/flock-core/src/).
#pragma once
#include <cstddef>
#include <cstdint>

#include "../../include/flock_gpu.h"

namespace flk {

int set_error(int code, const char* fmt, ...);
int check_enabled();
flk_f128 host_mul(flk_f128 a, flk_f128 b);
flk_f128 host_inv(flk_f128 a);

// phi8: GF(2^8) -> GF(2^128), fixed by phi8(0x02) (CF/field/phi8.rs:19-31,1047).
flk_f128 phi8_host(uint8_t v);

// L_i(z) for i < 2^k_skip on the nodes phi8(base .. base + 2^k_skip - 1):
// base 0 is the skip domain S (lagrange_weights_naive, CF/zerocheck/multilinear.rs:293),
// base 2^k_skip the extension domain Lambda (lagrange_weights_lambda_naive, :399-405).
// The node-only denominators are cached per (k_skip, base).  k_skip <= 7.
int lagrange_weights_nodes(size_t k_skip, flk_f128 z, size_t base, flk_f128* out);

// r[k_skip .. k_skip + 7) of the zerocheck: small_challenges_ghash (phi8 of 0xF7,
// 0x53, 0xB5) then medium_challenges_ghash (x^(2^i) / (1 + x^(2^i)), i < 4),
// CF/zerocheck/univariate_skip_optimized.rs:126-232.
void zc_fixed_challenges(flk_f128 out[7]);

inline flk_f128 fadd(flk_f128 a, flk_f128 b) { return flk_f128{a.lo ^ b.lo, a.hi ^ b.hi}; }

// flk_zc_round1 that also keeps the column sums of c in s_out (device, 128
// entries; NULL: not kept): S = fold_1b_rows(c as 128-bit elements, eq over
// r[7..m)), which is the ring-switch s_hat_v of a claim whose suffix point is
// r[7..m) over the same bits.
int zc_round1_keep_s(const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                     const flk_f128* r_outer, flk_f128* out_ab, flk_f128* out_c, flk_f128* s_out, void* stream);
// flk_pcs_combined_basis_and_target with some claims' s_hat_v already on the
// device: known[k] (128 entries) is taken as claim k's s_hat_v instead of
// folding the witness again (NULL entries, or known = NULL, fold as usual).
int pcs_combined_basis_and_target_known(flk_fs* fs, const flk_f128* packed_witness, size_t log_len,
                                        const flk_f128* x_outers, size_t n_claims, const flk_f128* const* known,
                                        flk_f128* s_hat_v_out, flk_f128* claims_out, flk_f128* gammas_out,
                                        flk_f128* b_combined, flk_f128* target_out, void* stream);

// Device arenas (runtime.hip): one device block carved by a best-fit free list.
// Inside an ArenaScope the calling thread's stream-ordered scratch (the library's
// temporaries and flk_alloc_async) comes from the arena, with no HIP call per
// allocation; a block freed there is handed out again at once, which is safe
// because everything in the scope is ordered on one stream.  Outside a scope,
// or when the arena is full, the stream-ordered pool serves the allocation.
class Arena;
class ArenaScope {
 public:
  explicit ArenaScope(Arena* a);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;

 private:
  Arena* prev_;
};
// A new arena of `bytes` (nullptr and the error recorded when the device
// allocation fails), its release (no block may be in use), its statistics.
Arena* arena_new(size_t bytes, int* status);
// An arena over `bytes` carved from `parent`, given back to the parent by
// arena_delete (nullptr and FLK_E_NOMEM when the parent has no room).
Arena* arena_view_new(Arena* parent, size_t bytes, int* status);
void arena_delete(Arena* a);
void arena_stats(const Arena* a, size_t* capacity, size_t* in_use, size_t* peak, size_t* fallbacks);

// The BLAKE3 prover of a setup (prover.cpp): the device bytes one prove draws
// from its arena, the setup's m, and a prove whose scratch comes from `arena`
// (the stream-ordered pool when null); same arguments as flk_blake3_prove.
size_t blake3_arena_bytes(const flk_blake3_setup* s);
size_t blake3_setup_m(const flk_blake3_setup* s);
int blake3_prove_in(const flk_blake3_setup* s, Arena* arena, const flk_compression* blocks, size_t n_blocks,
                    const uint8_t* domain, size_t domain_len, uint8_t* proof_out, size_t cap, size_t* proof_len,
                    double* stage_seconds, void* stream);

}  // namespace flk
