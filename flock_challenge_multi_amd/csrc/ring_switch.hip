// Ring-switch proving and the PCS claim combination (CF = crates/flock-core/src/):
//
//   flk_rs_prove_batched                ring_switch::prove_batched (CF/pcs/ring_switch.rs:3727-3735;
//                                       per-claim tail and gamma sampling :4310-4383,4620-4668)
//   flk_pcs_combined_basis_and_target   compute_combined_basis_and_target (CF/pcs.rs:778-870), for
//                                       ring-switched claims
//
// The large work is on the device: each claim's suffix tensor eq(x_outer[1..])
// (flk_build_eq; from 2^20 elements on only its two split factors), its
// s_hat_v = fold_1b_rows(z, suffix), and one fused pass that writes the
// combined basis
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

// The same sum over the split suffix tensor suffix_k[i] = lo_k[i mod B] * hi_k[i / B],
// B = 2^CS_LO_LOG (build_eq_split, CF/pcs/ring_switch.rs:331-357).  For a fixed
// e = hi_k[j], v -> fold(v * e) is GF(2)-linear in the bits of v, so block j
// folds lo_k against the 128-entry table T[c] = fold(x^c * e) (the column
// composition of compose_block_cols, :2130-2160): no suffix tensor is written or
// read, lo_k (64 KiB) stays in L2.  Each lane owns 16 elements of the block.
constexpr uint32_t CS_LO_LOG = 12;
constexpr uint32_t CS_PER_LANE = (1u << CS_LO_LOG) / RC_WG;
constexpr size_t CS_SPLIT_MIN_LOG = 20;

__device__ __forceinline__ uint4 fold_nib(uint4 v, const uint4* t) {
  const uint4 r0 = nibtab8_sum(v.x, t), r1 = nibtab8_sum(v.y, t + 128);
  const uint4 r2 = nibtab8_sum(v.z, t + 256), r3 = nibtab8_sum(v.w, t + 384);
  return make_uint4(xor3(r0.x, r1.x, r2.x) ^ r3.x, xor3(r0.y, r1.y, r2.y) ^ r3.y, xor3(r0.z, r1.z, r2.z) ^ r3.z,
                    xor3(r0.w, r1.w, r2.w) ^ r3.w);
}

__global__ void __launch_bounds__(RC_WG) rs_combine_split_kernel(const uint4* __restrict__ lo, const uint4* __restrict__ hi,
                                                                uint64_t n_hi, uint32_t n_claims,
                                                                const uint4* __restrict__ eqg, uint32_t accumulate,
                                                                uint4* __restrict__ out) {
  __shared__ uint4 gtab[RC_MAXC * 512];  // nibble tables of gamma_k eq(r''_k)
  __shared__ uint4 col[128];             // T[c] = fold(x^c * e) of the current (claim, block)
  __shared__ uint4 tnib[512];            // nibble tables of T
  const uint32_t t = threadIdx.x;
  for (uint32_t e = t; e < n_claims * 512u; e += RC_WG) {
    const uint32_t k = e >> 9, np = (e >> 4) & 31u, v = e & 15u;
    uint4 acc = f_zero();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((v >> j) & 1u) acc = f_add(acc, eqg[128 * k + 4 * np + j]);
    gtab[e] = acc;
  }
  __syncthreads();
  const uint64_t B = uint64_t(1) << CS_LO_LOG;
  for (uint64_t blk = blockIdx.x; blk < n_hi; blk += gridDim.x) {
    uint4 acc[CS_PER_LANE];
#pragma unroll
    for (uint32_t q = 0; q < CS_PER_LANE; ++q) acc[q] = accumulate ? out[blk * B + t + RC_WG * q] : f_zero();
    for (uint32_t k = 0; k < n_claims; ++k) {
      if (t < 128) {
        uint4 v = hi[k * n_hi + blk];  // x^t * e, in steps of at most 31 doublings
        for (uint32_t s = t; s; ) {
          const uint32_t d = s < 31u ? s : 31u;
          v = f_mul_xs(v, d);
          s -= d;
        }
        col[t] = fold_nib(v, gtab + 512 * k);
      }
      __syncthreads();
      for (uint32_t e = t; e < 512u; e += RC_WG) {
        const uint32_t np = e >> 4, v = e & 15u;
        uint4 a = f_zero();
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if ((v >> j) & 1u) a = f_add(a, col[4 * np + j]);
        tnib[e] = a;
      }
      __syncthreads();
      const uint4* lk = lo + k * B;
#pragma unroll
      for (uint32_t q = 0; q < CS_PER_LANE; ++q) acc[q] = f_add(acc[q], fold_nib(lk[t + RC_WG * q], tnib));
      __syncthreads();
    }
#pragma unroll
    for (uint32_t q = 0; q < CS_PER_LANE; ++q) out[blk * B + t + RC_WG * q] = acc[q];
  }
}

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
                  hipStream_t st, const flk_f128* const* known = nullptr) {
  const size_t n = size_t(1) << log_len, d = log_len + 1;  // x_outer has log_len + 1 coordinates
  // From 2^20 elements on (>= 256 blocks of 2^12, one work-group each), the
  // suffix tensor stays split: per claim eq_lo (2^12 elements) and eq_hi
  // (2^(log_len - 12)) instead of 2^log_len.
  const bool split = log_len >= CS_SPLIT_MIN_LOG;
  const size_t hi_log = split ? log_len - CS_LO_LOG : 0, n_hi = size_t(1) << hi_log;
  const size_t per_claim = split ? (size_t(1) << CS_LO_LOG) + n_hi : n;

  // Workspace: every claim's suffix tensor (or its two factors), its s_hat_v and
  // its gamma-scaled eq(r'').
  uint4* ws = nullptr;
  const size_t ws_elems = n_claims * (per_claim + 256);
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&ws), ws_elems * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "rs_prove_batched: %zu-byte workspace allocation failed", ws_elems * sizeof(uint4));
  }
  uint4* sfx = ws;  // dense: claim k at k * n; split: lo_k at k * B, then every hi_k at n_claims * B + k * n_hi
  uint4* shv = ws + n_claims * per_claim;
  uint4* eqg_dev = shv + n_claims * 128;
  std::vector<flk_f128> eqg(n_claims * 128);
  int rc = FLK_OK;
  auto finish = [&](int status) {
    hipError_t fe = flk::stream_free(ws, st);
    if (status == FLK_OK && fe != hipSuccess) return device_failure(fe, "rs_prove_batched");
    if (fe != hipSuccess) (void)hipGetLastError();
    return status;
  };

  // 1. Suffix tensors and s_hat_v (independent of the transcript); a known
  // s_hat_v is copied instead of folding the witness again.
  const size_t B = size_t(1) << CS_LO_LOG;
  uint4* his = sfx + n_claims * B;
  for (size_t k = 0; k < n_claims && rc == FLK_OK; ++k) {
    flk_f128* out128 = reinterpret_cast<flk_f128*>(shv + 128 * k);
    const flk_f128* kn = known ? known[k] : nullptr;
    if (split) {
      flk_f128* lo = reinterpret_cast<flk_f128*>(sfx + k * B);
      flk_f128* hi = reinterpret_cast<flk_f128*>(his + k * n_hi);
      rc = flk_build_eq(x_outers + k * d + 1, CS_LO_LOG, lo, st);
      if (rc == FLK_OK) rc = flk_build_eq(x_outers + k * d + 1 + CS_LO_LOG, hi_log, hi, st);
      if (rc == FLK_OK && !kn) rc = rs_fold_1b_rows_split(z, lo, CS_LO_LOG, hi, hi_log, out128, st);
    } else {
      flk_f128* sk = reinterpret_cast<flk_f128*>(sfx + k * n);
      rc = flk_build_eq(x_outers + k * d + 1, log_len, sk, st);
      if (rc == FLK_OK && !kn) rc = flk_rs_fold_1b_rows(z, sk, n, out128, st);
    }
    if (rc == FLK_OK && kn) {
      e = hipMemcpyAsync(out128, kn, 128 * sizeof(flk_f128), hipMemcpyDeviceToDevice, st);
      if (e != hipSuccess) rc = device_failure(e, "rs_prove_batched: known s_hat_v copy");
    }
  }
  if (rc != FLK_OK) return finish(rc);
  e = read_to_host(s_hat_v_out, shv, n_claims * 128 * sizeof(uint4), st);
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
  e = upload_staged(eqg_dev, eqg.data(), eqg.size() * sizeof(flk_f128), st);
  if (e != hipSuccess) return finish(device_failure(e, "rs_prove_batched: eq upload"));
  const unsigned grid = split ? (unsigned)std::min<uint64_t>(n_hi, 2048)
                             : (unsigned)std::min<uint64_t>((n + RC_WG - 1) / RC_WG, 1024);
  for (size_t k0 = 0; k0 < n_claims; k0 += RC_MAXC) {
    const uint32_t nc = (uint32_t)std::min<size_t>(RC_MAXC, n_claims - k0);
    if (split) {
      hipLaunchKernelGGL(rs_combine_split_kernel, dim3(grid), dim3(RC_WG), 0, st, sfx + k0 * B, his + k0 * n_hi,
                         (uint64_t)n_hi, nc, eqg_dev + 128 * k0, (uint32_t)(k0 != 0),
                         reinterpret_cast<uint4*>(b_combined));
      FLK_LAUNCH_CHECK("rs_combine_split_kernel");
    } else {
      hipLaunchKernelGGL(rs_combine_kernel, dim3(grid), dim3(RC_WG), 0, st, sfx + k0 * n, (uint64_t)n, nc,
                         eqg_dev + 128 * k0, (uint32_t)(k0 != 0), reinterpret_cast<uint4*>(b_combined));
      FLK_LAUNCH_CHECK("rs_combine_kernel");
    }
  }
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
  return flk::pcs_combined_basis_and_target_known(fs, packed_witness, log_len, x_outers, n_claims, nullptr,
                                                  s_hat_v_out, claims_out, gammas_out, b_combined, target_out,
                                                  stream);
}

}  // extern "C"

int flk::pcs_combined_basis_and_target_known(flk_fs* fs, const flk_f128* packed_witness, size_t log_len,
                                             const flk_f128* x_outers, size_t n_claims,
                                             const flk_f128* const* known, flk_f128* s_hat_v_out,
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
                       b_combined, static_cast<hipStream_t>(stream), known);
  if (rc != FLK_OK) return rc;
  flk_f128 t{0, 0};
  for (size_t k = 0; k < n_claims; ++k) t = fadd(t, host_mul(gammas_out[k], claims_out[k]));
  *target_out = t;
  return FLK_OK;
}
