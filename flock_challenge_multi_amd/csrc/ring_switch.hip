// Ring-switch proving and the PCS claim combination (CF = crates/flock-core/src/):
//
//   flk_rs_prove_batched                ring_switch::prove_batched (CF/pcs/ring_switch.rs:3727-3735;
//                                       per-claim tail and gamma sampling :4310-4383,4620-4668)
//   flk_pcs_combined_basis_and_target   compute_combined_basis_and_target (CF/pcs.rs:778-870), for
//                                       ring-switched claims
//
// The large work is on the device: each claim's suffix tensor eq(x_outer[1..])
// (flk_build_eq), its s_hat_v = fold_1b_rows(z, suffix) (flk_rs_fold_1b_rows),
// and one fused pass that writes the combined basis
//   b[i] = sum_k sum_b bit_b(suffix_k[i]) * gamma_k eq(r''_k)[b]
// reading every claim's suffix tensor once (nibble tables of gamma_k eq(r''_k)
// per claim in LDS).  The reference bakes gamma_k into eq(r''_k) before its
// fold (:4380-4383) so the combination is a plain sum; the fused pass forms that
// sum without writing any per-claim rs_eq_ind.  The 128-element steps (the
// transcript, eq(r''), the tensor-algebra transpose and the claim's inner
// product, CF/pcs/ring_switch.rs:1646-1680,1757-1950) run on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {
namespace {

constexpr int RC_WG = 256;
constexpr uint32_t RC_MAXC = 4;  // claims per fused pass (8 KiB of nibble tables each)
constexpr size_t LOG_PACKING = 7;

// out[i] (+)= sum_{k < n_claims} sum_b bit_b(sfx[k * n + i]) * eqg[128 k + b].
__global__ void __launch_bounds__(RC_WG) rs_combine_kernel(const uint4* __restrict__ sfx, uint64_t n,
                                                          uint32_t n_claims, const uint4* __restrict__ eqg,
                                                          uint32_t accumulate, uint4* __restrict__ out) {
  __shared__ uint4 nibt[RC_MAXC * 32 * 16];
  for (uint32_t e = threadIdx.x; e < n_claims * 512u; e += RC_WG) {
    const uint32_t k = e >> 9, np = (e >> 4) & 31u, v = e & 15u;  // claim, nibble position, nibble value
    uint4 acc = f_zero();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((v >> j) & 1u) acc = f_add(acc, eqg[128 * k + 4 * np + j]);
    nibt[e] = acc;
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * RC_WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RC_WG) {
    uint4 acc = accumulate ? out[i] : f_zero();
    for (uint32_t k = 0; k < n_claims; ++k) {
      const uint4 s = sfx[(uint64_t)k * n + i];
      const uint4* t = nibt + 512 * k;
      const uint4 r0 = nibtab8_sum(s.x, t), r1 = nibtab8_sum(s.y, t + 128);
      const uint4 r2 = nibtab8_sum(s.z, t + 256), r3 = nibtab8_sum(s.w, t + 384);
      acc = make_uint4(acc.x ^ xor3(r0.x, r1.x, r2.x) ^ r3.x, acc.y ^ xor3(r0.y, r1.y, r2.y) ^ r3.y,
                       acc.z ^ xor3(r0.z, r1.z, r2.z) ^ r3.z, acc.w ^ xor3(r0.w, r1.w, r2.w) ^ r3.w);
    }
    out[i] = acc;
  }
}

inline flk_f128 fadd(flk_f128 a, flk_f128 b) { return flk_f128{a.lo ^ b.lo, a.hi ^ b.hi}; }

// build_eq (CF/zerocheck/univariate_skip.rs:31-73) on the host: LSB-first table of 2^d entries.
std::vector<flk_f128> host_eq(const flk_f128* r, size_t d) {
  std::vector<flk_f128> t(size_t(1) << d, flk_f128{0, 0});
  t[0] = flk_f128{1, 0};
  for (size_t i = 0; i < d; ++i)
    for (size_t x = (size_t(1) << i); x-- > 0;) {
      const flk_f128 hi = host_mul(t[x], r[i]);
      t[x | (size_t(1) << i)] = hi;
      t[x] = fadd(t[x], hi);
    }
  return t;
}

// sumcheck_claim = <tensor_algebra_transpose(s_hat_v), eq_r> (CF/pcs/ring_switch.rs:1700-1703,
// 1916-1940): s_hat_u[b] has bit i set iff bit b of s_hat_v[i] is set.
flk_f128 claim_from_s_hat_v(const flk_f128* s_hat_v, const flk_f128* eq_r) {
  flk_f128 acc{0, 0};
  for (unsigned b = 0; b < 128; ++b) {
    flk_f128 u{0, 0};
    for (unsigned i = 0; i < 128; ++i) {
      const uint64_t w = b < 64 ? s_hat_v[i].lo : s_hat_v[i].hi;
      if ((w >> (b & 63)) & 1u) {
        if (i < 64) u.lo |= uint64_t(1) << i;
        else u.hi |= uint64_t(1) << (i - 64);
      }
    }
    acc = fadd(acc, host_mul(u, eq_r[b]));
  }
  return acc;
}

const uint8_t RS_LABEL[] = "flock-ring-switch-v0";
const uint8_t PCS_LABEL[] = "flock-pcs-open-batch-v0";

// Shape checks, made before the transcript is touched.
int validate(const char* what, flk_fs* fs, const flk_f128* z, size_t log_len, const flk_f128* x_outers,
             size_t n_claims, const flk_f128* s_hat_v_out, const flk_f128* claims_out, const flk_f128* gammas_out) {
  if (!fs || !z || !x_outers || !s_hat_v_out || !claims_out || !gammas_out)
    return set_error(FLK_E_SHAPE, "%s: null pointer", what);
  if (n_claims == 0) return set_error(FLK_E_SHAPE, "%s: no claims", what);
  if (n_claims > 64) return set_error(FLK_E_SHAPE, "%s: %zu claims (at most 64)", what, n_claims);
  if (log_len > 34) return set_error(FLK_E_SHAPE, "%s: log_len=%zu too large", what, log_len);
  return FLK_OK;
}

int prove_batched(flk_fs* fs, const flk_f128* z, size_t log_len, const flk_f128* x_outers, size_t n_claims,
                  flk_f128* s_hat_v_out, flk_f128* claims_out, flk_f128* gammas_out, flk_f128* b_combined,
                  hipStream_t st) {
  const size_t n = size_t(1) << log_len, d = log_len + 1;  // x_outer has log_len + 1 coordinates

  // Workspace: every claim's suffix tensor, its s_hat_v and its gamma-scaled eq(r'').
  uint4* ws = nullptr;
  const size_t ws_elems = n_claims * (n + 256);
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&ws), ws_elems * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "rs_prove_batched: %zu-byte workspace allocation failed", ws_elems * sizeof(uint4));
  }
  uint4* sfx = ws;
  uint4* shv = ws + n_claims * n;
  uint4* eqg_dev = shv + n_claims * 128;
  std::vector<flk_f128> eqg(n_claims * 128);
  int rc = FLK_OK;
  auto finish = [&](int status) {
    hipError_t fe = hipFreeAsync(ws, st);
    if (status == FLK_OK && fe != hipSuccess) return device_failure(fe, "rs_prove_batched");
    if (fe != hipSuccess) (void)hipGetLastError();
    return status;
  };

  // 1. Suffix tensors and s_hat_v (independent of the transcript).
  for (size_t k = 0; k < n_claims && rc == FLK_OK; ++k) {
    flk_f128* sk = reinterpret_cast<flk_f128*>(sfx + k * n);
    rc = flk_build_eq(x_outers + k * d + 1, log_len, sk, st);
    if (rc == FLK_OK) rc = flk_rs_fold_1b_rows(z, sk, n, reinterpret_cast<flk_f128*>(shv + 128 * k), st);
  }
  if (rc != FLK_OK) return finish(rc);
  e = hipMemcpyAsync(s_hat_v_out, shv, n_claims * 128 * sizeof(uint4), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return finish(device_failure(e, "rs_prove_batched: s_hat_v readback"));

  // 2. Per claim: label, observe s_hat_v, sample r'' (7), the BaseFold claim.
  std::vector<std::vector<flk_f128>> eq_r(n_claims);
  for (size_t k = 0; k < n_claims && rc == FLK_OK; ++k) {
    flk_f128 r[LOG_PACKING];
    rc = flk_fs_observe_label(fs, RS_LABEL, sizeof(RS_LABEL) - 1);
    if (rc == FLK_OK) rc = flk_fs_observe_f128_slice(fs, s_hat_v_out + 128 * k, 128);
    if (rc == FLK_OK) rc = flk_fs_sample_f128_vec(fs, LOG_PACKING, r);
    if (rc != FLK_OK) break;
    eq_r[k] = host_eq(r, LOG_PACKING);
    claims_out[k] = claim_from_s_hat_v(s_hat_v_out + 128 * k, eq_r[k].data());
  }
  // 3. gamma_k after every claim is observed; gamma_k baked into eq(r''_k).
  for (size_t k = 0; k < n_claims && rc == FLK_OK; ++k) rc = flk_fs_sample_f128(fs, gammas_out + k);
  if (rc != FLK_OK || !b_combined) return finish(rc);
  for (size_t k = 0; k < n_claims; ++k)
    for (unsigned b = 0; b < 128; ++b) eqg[128 * k + b] = host_mul(gammas_out[k], eq_r[k][b]);

  // 4. b = sum_k fold_b128_elems(suffix_k, gamma_k eq(r''_k)), RC_MAXC claims per pass.
  e = hipMemcpyAsync(eqg_dev, eqg.data(), eqg.size() * sizeof(flk_f128), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return finish(device_failure(e, "rs_prove_batched: eq upload"));
  const unsigned grid = (unsigned)std::min<uint64_t>((n + RC_WG - 1) / RC_WG, 1024);
  for (size_t k0 = 0; k0 < n_claims; k0 += RC_MAXC) {
    const uint32_t nc = (uint32_t)std::min<size_t>(RC_MAXC, n_claims - k0);
    hipLaunchKernelGGL(rs_combine_kernel, dim3(grid), dim3(RC_WG), 0, st, sfx + k0 * n, (uint64_t)n, nc,
                       eqg_dev + 128 * k0, (uint32_t)(k0 != 0), reinterpret_cast<uint4*>(b_combined));
    FLK_LAUNCH_CHECK("rs_combine_kernel");
  }
  // the host eq table must outlive its upload
  e = hipStreamSynchronize(st);
  if (e != hipSuccess) return finish(device_failure(e, "rs_prove_batched"));
  return finish(FLK_OK);
}

}  // namespace
}  // namespace flk

using namespace flk;

extern "C" {

int flk_rs_prove_batched(flk_fs* fs, const flk_f128* packed_witness, size_t log_len, const flk_f128* x_outers,
                         size_t n_claims, flk_f128* s_hat_v_out, flk_f128* claims_out, flk_f128* gammas_out,
                         flk_f128* b_combined, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if ((rc = validate("rs_prove_batched", fs, packed_witness, log_len, x_outers, n_claims, s_hat_v_out, claims_out,
                     gammas_out)))
    return rc;
  return prove_batched(fs, packed_witness, log_len, x_outers, n_claims, s_hat_v_out, claims_out, gammas_out,
                       b_combined, static_cast<hipStream_t>(stream));
}

int flk_pcs_combined_basis_and_target(flk_fs* fs, const flk_f128* packed_witness, size_t log_len,
                                      const flk_f128* x_outers, size_t n_claims, flk_f128* s_hat_v_out,
                                      flk_f128* claims_out, flk_f128* gammas_out, flk_f128* b_combined,
                                      flk_f128* target_out, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if ((rc = validate("pcs_combined_basis_and_target", fs, packed_witness, log_len, x_outers, n_claims, s_hat_v_out,
                     claims_out, gammas_out)))
    return rc;
  if (!b_combined || !target_out) return set_error(FLK_E_SHAPE, "pcs_combined_basis_and_target: null pointer");
  rc = flk_fs_observe_label(fs, PCS_LABEL, sizeof(PCS_LABEL) - 1);
  if (rc == FLK_OK)
    rc = prove_batched(fs, packed_witness, log_len, x_outers, n_claims, s_hat_v_out, claims_out, gammas_out,
                       b_combined, static_cast<hipStream_t>(stream));
  if (rc != FLK_OK) return rc;
  flk_f128 t{0, 0};
  for (size_t k = 0; k < n_claims; ++k) t = fadd(t, host_mul(gammas_out[k], claims_out[k]));
  *target_out = t;
  return FLK_OK;
}

}  // extern "C"
