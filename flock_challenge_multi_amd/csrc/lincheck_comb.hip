// The lincheck circuit of the BLAKE3 R1CS on the device (CF = crates/flock-core/src/):
//
//   flk_blake3_r1cs_new / flk_r1cs_*   build_block_r1cs (CP/r1cs_hashes/blake3.rs:1531-1545)
//                                      with its CSC circuit (CF/lincheck.rs:1505-1560) in HBM
//                                      and statement_digest (CF/r1cs.rs:334-385)
//   flk_lincheck_comb                  comb_vec of lincheck's prover and verifier
//                                      (CF/lincheck.rs:5590-5660, 5945-5965):
//                                      fold_alpha_batched(alpha, build_quirky_eq_table(
//                                      z_skip, x_inner_rest, 6)) + beta at the pinned column
//
// comb[c] = alpha * sum_{r in colA(c)} eq[r] + sum_{r in colB(c)} eq[r]; the alpha
// product is moved into the table (eqa[r] = alpha * eq[r] = (alpha * L)[r mod 64] *
// eq_rest[r >> 6]), so the gather is pure XOR.  One wave per column: its lanes
// stride the column's 16-bit row list (21 M entries over both matrices, ~42 MB
// streamed once) and gather eq / eqa from a 2 x 256 KiB table that stays in L2;
// a wave XOR-reduction finishes the column.  Columns are independent, so the
// result does not depend on the schedule.
#include <hip/hip_runtime.h>

#include <mutex>
#include <new>

#include "gf128.hpp"
#include "internal.hpp"
#include "r1cs.hpp"

namespace {

using flk::f_add;
using flk::f_mul_x;
using flk::f_zero;
using flk::r1cs::K;
using flk::r1cs::K_LOG;
using flk::r1cs::K_SKIP;

constexpr unsigned N_SKIP = 1u << K_SKIP;            // 64 Lagrange nodes
constexpr unsigned N_REST = K_LOG - K_SKIP;          // 8 multilinear inner coordinates

// Bit-serial product (128 shift/XOR steps): only set-up tables use it.
__device__ uint4 f_mul_serial(uint4 a, uint4 b) {
  uint4 acc = f_zero();
  const uint32_t aw[4] = {a.x, a.y, a.z, a.w};
#pragma unroll 1
  for (int w = 0; w < 4; ++w)
#pragma unroll 1
    for (int i = 0; i < 32; ++i) {
      if ((aw[w] >> i) & 1u) acc = f_add(acc, b);
      b = f_mul_x(b);
    }
  return acc;
}

struct QuirkyArgs {
  uint4 lag[N_SKIP];    // L_i(z_skip) on the phi8 nodes 0..63
  uint4 alag[N_SKIP];   // alpha * L_i(z_skip)
  uint4 xr[N_REST];     // x_inner_rest
};

// eq[i] = L[i mod 64] * eq(x_inner_rest, i >> 6) (build_quirky_eq_table, CF/lincheck.rs:4653-4665),
// eqa[i] = alpha * eq[i].
__global__ void lc_quirky_eq_kernel(QuirkyArgs args, uint4* __restrict__ eq, uint4* __restrict__ eqa) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K) return;
  const unsigned rest = i >> K_SKIP;
  uint4 e = make_uint4(1u, 0u, 0u, 0u);
#pragma unroll 1
  for (unsigned j = 0; j < N_REST; ++j) {
    uint4 f = args.xr[j];
    if (!((rest >> j) & 1u)) f.x ^= 1u;  // 1 + x_j
    e = j ? f_mul_serial(e, f) : f;
  }
  eq[i] = f_mul_serial(args.lag[i & (N_SKIP - 1)], e);
  eqa[i] = f_mul_serial(args.alag[i & (N_SKIP - 1)], e);
}

__device__ __forceinline__ uint4 wave_xor(uint4 v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1)
    v = f_add(v, make_uint4(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m), __shfl_xor(v.w, m)));
  return v;
}

constexpr unsigned COMB_WAVES = 4;  // columns per 256-lane work-group

__global__ __launch_bounds__(64 * COMB_WAVES) void lc_comb_gather_kernel(
    const uint32_t* __restrict__ a_ptr, const uint16_t* __restrict__ a_rows, const uint32_t* __restrict__ b_ptr,
    const uint16_t* __restrict__ b_rows, const uint4* __restrict__ eq, const uint4* __restrict__ eqa,
    unsigned pin_col, uint4 beta, uint4* __restrict__ comb) {
  const unsigned lane = threadIdx.x & 63u;
  const unsigned col = blockIdx.x * COMB_WAVES + (threadIdx.x >> 6);
  if (col >= K) return;
  uint4 acc = f_zero();
  for (uint32_t j = a_ptr[col] + lane, e = a_ptr[col + 1]; j < e; j += 64u) acc = f_add(acc, eqa[a_rows[j]]);
  for (uint32_t j = b_ptr[col] + lane, e = b_ptr[col + 1]; j < e; j += 64u) acc = f_add(acc, eq[b_rows[j]]);
  acc = wave_xor(acc);
  if (lane == 0) comb[col] = col == pin_col ? f_add(acc, beta) : acc;
}

// The CSC circuit in HBM, one copy per device shared by every R1CS handle.
struct DeviceCsc {
  uint32_t *a_ptr = nullptr, *b_ptr = nullptr;
  uint16_t *a_rows = nullptr, *b_rows = nullptr;
};

int device_csc(DeviceCsc* out) {
  static std::mutex mu;
  static DeviceCsc per_dev[64];
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return flk::device_failure(e, "hipGetDevice");
  if (dev < 0 || dev >= 64) return flk::set_error(FLK_E_SHAPE, "r1cs: device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(mu);
  DeviceCsc& d = per_dev[dev];
  if (!d.a_ptr) {
    const flk::r1cs::Csc& c = flk::r1cs::blake3_csc();
    DeviceCsc n;
    const size_t pb = (K + 1) * sizeof(uint32_t);
    if ((e = hipMalloc(&n.a_ptr, pb)) != hipSuccess || (e = hipMalloc(&n.b_ptr, pb)) != hipSuccess ||
        (e = hipMalloc(&n.a_rows, c.a_rows.size() * 2)) != hipSuccess ||
        (e = hipMalloc(&n.b_rows, c.b_rows.size() * 2)) != hipSuccess) {
      (void)hipFree(n.a_ptr);
      (void)hipFree(n.b_ptr);
      (void)hipFree(n.a_rows);
      (void)hipFree(n.b_rows);
      return flk::set_error(FLK_E_NOMEM, "r1cs: device CSC allocation failed");
    }
    if ((e = hipMemcpy(n.a_ptr, c.a_ptr.data(), pb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(n.b_ptr, c.b_ptr.data(), pb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(n.a_rows, c.a_rows.data(), c.a_rows.size() * 2, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(n.b_rows, c.b_rows.data(), c.b_rows.size() * 2, hipMemcpyHostToDevice)) != hipSuccess)
      return flk::device_failure(e, "r1cs: CSC upload");
    d = n;
  }
  *out = d;
  return FLK_OK;
}

inline uint4 u4(flk_f128 v) { return make_uint4((uint32_t)v.lo, (uint32_t)(v.lo >> 32), (uint32_t)v.hi, (uint32_t)(v.hi >> 32)); }

}  // namespace

struct flk_r1cs {
  size_t m;
  DeviceCsc csc;
};

// Host Lagrange weights on the phi8 nodes 0..63 (declared in flock_gpu.h).
extern "C" int flk_zc_lagrange_weights(size_t k_skip, flk_f128 z, flk_f128* out);

extern "C" {

int flk_blake3_r1cs_new(size_t n_blocks_log, flk_r1cs** out) {
  if (int rc = flk::check_enabled()) return rc;
  if (!out) return flk::set_error(FLK_E_SHAPE, "blake3_r1cs_new: null pointer");
  if (n_blocks_log < 3 || n_blocks_log > 26)
    return flk::set_error(FLK_E_SHAPE, "blake3_r1cs_new: n_blocks_log=%zu outside 3..26", n_blocks_log);
  DeviceCsc d;
  if (int rc = device_csc(&d)) return rc;
  flk_r1cs* r = new (std::nothrow) flk_r1cs{K_LOG + n_blocks_log, d};
  if (!r) return flk::set_error(FLK_E_NOMEM, "blake3_r1cs_new: out of host memory");
  *out = r;
  return FLK_OK;
}

void flk_r1cs_free(flk_r1cs* r) { delete r; }

int flk_r1cs_shape(const flk_r1cs* r, size_t* m, size_t* k_log, size_t* k_skip, size_t* nnz_a, size_t* nnz_b) {
  if (!r) return flk::set_error(FLK_E_SHAPE, "r1cs_shape: null pointer");
  const flk::r1cs::Matrices& mt = flk::r1cs::blake3_matrices();
  if (m) *m = r->m;
  if (k_log) *k_log = K_LOG;
  if (k_skip) *k_skip = K_SKIP;
  if (nnz_a) *nnz_a = mt.a_cols.size();
  if (nnz_b) *nnz_b = mt.b_cols.size();
  return FLK_OK;
}

int flk_r1cs_statement_digest(const flk_r1cs* r, uint8_t out[32]) {
  if (!r || !out) return flk::set_error(FLK_E_SHAPE, "r1cs_statement_digest: null pointer");
  flk::r1cs::statement_digest(r->m, out);
  return FLK_OK;
}

int flk_lincheck_comb(const flk_r1cs* r, flk_f128 z_skip, const flk_f128* x_inner_rest, size_t n_rest,
                      flk_f128 alpha, flk_f128 beta, flk_f128* comb_out, void* stream) {
  if (int rc = flk::check_enabled()) return rc;
  if (!r || !x_inner_rest || !comb_out) return flk::set_error(FLK_E_SHAPE, "lincheck_comb: null pointer");
  if (n_rest != N_REST) return flk::set_error(FLK_E_SHAPE, "lincheck_comb: %zu inner-rest coordinates, need %u", n_rest, N_REST);
  hipStream_t st = static_cast<hipStream_t>(stream);
  QuirkyArgs args;
  flk_f128 lag[N_SKIP];
  if (int rc = flk_zc_lagrange_weights(K_SKIP, z_skip, lag)) return rc;
  for (unsigned i = 0; i < N_SKIP; ++i) {
    args.lag[i] = u4(lag[i]);
    args.alag[i] = u4(flk::host_mul(alpha, lag[i]));
  }
  for (unsigned j = 0; j < N_REST; ++j) args.xr[j] = u4(x_inner_rest[j]);
  uint4* tabs = nullptr;
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&tabs), 2 * K * sizeof(uint4), st);
  if (e != hipSuccess) return flk::set_error(FLK_E_NOMEM, "lincheck_comb: scratch allocation failed");
  hipLaunchKernelGGL(lc_quirky_eq_kernel, dim3(K / 256), dim3(256), 0, st, args, tabs, tabs + K);
  FLK_LAUNCH_CHECK("lc_quirky_eq_kernel");
  hipLaunchKernelGGL(lc_comb_gather_kernel, dim3(K / COMB_WAVES), dim3(64 * COMB_WAVES), 0, st, r->csc.a_ptr,
                     r->csc.a_rows, r->csc.b_ptr, r->csc.b_rows, tabs, tabs + K, (unsigned)flk::r1cs::CONST_POS,
                     u4(beta), reinterpret_cast<uint4*>(comb_out));
  FLK_LAUNCH_CHECK("lc_comb_gather_kernel");
  if ((e = flk::stream_free(tabs, st)) != hipSuccess) return flk::device_failure(e, "lincheck_comb: hipFreeAsync");
  return FLK_OK;
}

}  // extern "C"
