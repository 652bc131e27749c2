// The BLAKE3 compression R1CS (host side): the per-block base matrices A_0 and
// B_0, their column-major form for the lincheck fold, and the statement digest
// (CP = crates/flock-prover/src/, CF = crates/flock-core/src/).
#pragma once
#include <cstddef>
#include <cstdint>
#include <vector>

namespace flk {
namespace r1cs {

// Block layout (CP/r1cs_hashes/blake3.rs:111-241).
constexpr size_t K_LOG = 14;
constexpr size_t K = size_t(1) << K_LOG;
constexpr size_t K_SKIP = 6;
constexpr size_t CONST_POS = 512;       // the constant-one wire, pinned by lincheck
constexpr size_t USEFUL_BITS = 15409;   // rows [USEFUL_BITS, K) are zero padding

// A_0, B_0 (build_matrices, blake3.rs:483-632) in CSR form: row r's columns are
// cols[ptr[r] .. ptr[r+1]), strictly ascending (xor_dedup sorts and cancels,
// CP/r1cs_hashes/common.rs:437-453).  C_0 is the identity and not stored.
struct Matrices {
  std::vector<uint32_t> a_ptr, a_cols, b_ptr, b_cols;
};
// Built once per process (~0.1 s).
const Matrices& blake3_matrices();

// The same matrices by column (CscCircuit, CF/lincheck.rs:1505-1560): column c's
// rows are rows[ptr[c] .. ptr[c+1]) (ascending), 16-bit since K = 2^14.
struct Csc {
  std::vector<uint32_t> a_ptr, b_ptr;
  std::vector<uint16_t> a_rows, b_rows;
};
const Csc& blake3_csc();

// BlockR1cs::statement_digest (CF/r1cs.rs:334-385) of the BLAKE3 R1CS with
// m = K_LOG + n_blocks_log, RowMajor layout: BLAKE3 over
// "flock-r1cs-stmt-v1" || u64le(m) || u64le(k_log) || u64le(k_skip) || 0x00 ||
// absorb_matrix(A_0) || absorb_matrix(B_0) || absorb_matrix(I_K), where
// absorb_matrix = u64le(rows) || u64le(cols) || per row u64le(len) || u64le(col)*.
// Cached per m; thread-safe.
void statement_digest(size_t m, uint8_t out[32]);

}  // namespace r1cs
}  // namespace flk
