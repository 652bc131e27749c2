// Ligerito basis induction on gfx950 (induce_sumcheck_poly,
// CF/pcs/ligerito.rs:1436-1548, with evaluate_scaled_basis_inplace :1237-1300
// and eval_sk_at_vks :1186-1205).
//
//   basis[j] = sum_q alpha_q * prod_{k in bits(j)} s_k(x_q) / s_k(v_k)
//   sum      = sum_q alpha_q * <row_q, eq(v_challenges)>
//
// Device schedule (products by one shared multiplier use a work-group comb table):
//   1. one lane per query walks the subspace-polynomial recurrence
//      s_0 = x, s_i = s_{i-1}^2 + s_{i-1}(v_{i-1}) s_{i-1}, scaled by 1/s_i(v_i);
//   2. per query, the low and high halves of the tensor expansion
//      lo_q[j] (2^s_lo entries) and hi_q[j] = alpha_q * (2^s_hi entries);
//   3. basis[(jh << s_lo) | jl] = sum_q lo_q[jl] * hi_q[jh]: a work-group owns one jh
//      and 256 jl, hi_q[jh] is its shared multiplier, lanes XOR-accumulate
//      unreduced products;
//   4. the enforced sum as a dot product of the opened rows with alpha_q eq_k
//      (both factors applied on the device).
// Every sum is an XOR, so this order equals the reference's per-thread partials.
#include <hip/hip_runtime.h>

#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

constexpr int IND_WG = 256;

__device__ __forceinline__ uint4 shfl_xor4_i(uint4 v, int m) {
  return make_uint4(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m), __shfl_xor(v.w, m));
}

// 1. s_k(x_q) / s_k(v_k) for k < L.  consts = [sks_vks (L+1) | inv_sks (L+1)].
__global__ void __launch_bounds__(64) induce_sks_kernel(const uint64_t* __restrict__ queries, uint32_t nq, uint32_t L,
                                                       const uint4* __restrict__ consts, uint4* __restrict__ S) {
  __shared__ uint4 tabs[16 * 64];
  const uint32_t q = blockIdx.x * 64 + threadIdx.x;
  const uint64_t x = q < nq ? queries[q] : 0ull;
  uint4 s = make_uint4((uint32_t)x, (uint32_t)(x >> 32), 0u, 0u);
  uint4* my = tabs + threadIdx.x;
  for (uint32_t i = 0; i < L; ++i) {
    if (i > 0) {
      // s <- s*s + sks_vks[i-1]*s = s * (s + sks_vks[i-1])
      table_store_private(tabs, 64, threadIdx.x, s);
      s = comb_mul_rt(f_add(s, consts[i - 1]), my, 64);
    }
    table_store_private(tabs, 64, threadIdx.x, s);
    const uint4 scaled = comb_mul_rt(consts[(L + 1) + i], my, 64);
    if (q < nq) S[(size_t)q * L + i] = scaled;
  }
}

// 2. Tensor halves of query q (one work-group per query and half).
//    out[q * 2^bits + j] = base * prod_{i in bits(j)} S[q*L + first + i].
__global__ void __launch_bounds__(IND_WG) induce_expand_kernel(const uint4* __restrict__ S, uint32_t L, uint32_t first,
                                                              uint32_t bits, const uint4* __restrict__ base_per_q,
                                                              uint4* __restrict__ out) {
  __shared__ uint4 tab[16];
  const uint32_t q = blockIdx.x;
  uint4* o = out + ((size_t)q << bits);
  if (threadIdx.x == 0) o[0] = base_per_q ? base_per_q[q] : make_uint4(1u, 0u, 0u, 0u);
  for (uint32_t i = 0; i < bits; ++i) {
    const uint32_t len = 1u << i;
    if (threadIdx.x < 16) tab[threadIdx.x] = comb_entry(S[(size_t)q * L + first + i], threadIdx.x);
    __threadfence_block();
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < len; j += IND_WG) o[len + j] = comb_mul<1>(o[j], tab);
    __threadfence_block();
    __syncthreads();
  }
}

// 3. basis[(jh << s_lo) | jl] = sum_q lo[q][jl] * hi[q][jh]; work-group (jh, y)
//    takes jl in [y * IND_WG * ITEMS, (y + 1) * IND_WG * ITEMS), so a large
//    basis runs as 2^s_hi x 2^s_lo / (IND_WG * ITEMS) work-groups (each builds the
//    nibble-position tables of its hi[.][jh]).
template <int ITEMS>
__global__ void __launch_bounds__(IND_WG) induce_basis_kernel(const uint4* __restrict__ lo, const uint4* __restrict__ hi,
                                                             uint32_t nq, uint32_t s_lo, uint32_t s_hi,
                                                             uint4* __restrict__ basis) {
  constexpr uint32_t QC = 8;  // queries per table build
  __shared__ uint4 tab[QC * 128];  // nibble-position tables of hi[q][jh], q in the chunk
  __shared__ uint4 nib[QC * 32];
  const uint32_t jh = blockIdx.x;
  const uint32_t n_lo = 1u << s_lo;
  const uint32_t jl0 = blockIdx.y * (IND_WG * ITEMS);
  uint32_t acc[ITEMS][7];
#pragma unroll
  for (int t = 0; t < ITEMS; ++t)
#pragma unroll
    for (int w = 0; w < 7; ++w) acc[t][w] = 0u;
  for (uint32_t q0 = 0; q0 < nq; q0 += QC) {
    const uint32_t nc = nq - q0 < QC ? nq - q0 : QC;
    __syncthreads();  // the previous chunk's lookups are done
    build_pos_tables(tab, nib, nc, [&](uint32_t c) { return hi[((size_t)(q0 + c) << s_hi) + jh]; });
    for (uint32_t c = 0; c < nc; ++c) {
#pragma unroll
      for (int t = 0; t < ITEMS; ++t) {
        const uint32_t jl = jl0 + threadIdx.x + t * IND_WG;
        if (jl < n_lo) combjr_acc(lo[((size_t)(q0 + c) << s_lo) + jl], tab + 128 * c, acc[t]);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < ITEMS; ++t) {
    const uint32_t jl = jl0 + threadIdx.x + t * IND_WG;
    if (jl < n_lo) {
      const uint32_t r[8] = {acc[t][0], acc[t][1], acc[t][2], acc[t][3], acc[t][4], acc[t][5], acc[t][6], 0u};
      basis[((size_t)jh << s_lo) + jl] = f_reduce_add(r, f_zero());
    }
  }
}

// 4. sum over rows[q * 2^log_row + k] * eqv[k] * aq[q], XOR-accumulated into *out
//    (out zeroed by the caller): two products per element on the lane's private
//    comb tables.
__global__ void __launch_bounds__(IND_WG) dot_kernel(const uint4* __restrict__ rows, const uint4* __restrict__ eqv,
                                                    const uint4* __restrict__ aq, uint32_t log_row, uint64_t n,
                                                    uint4* __restrict__ out) {
  __shared__ uint4 tabs[16 * IND_WG];
  __shared__ uint4 red[IND_WG / 64];
  U256 acc;
  u256_zero(acc);
  const uint64_t kmask = (uint64_t(1) << log_row) - 1;
  for (uint64_t i = (uint64_t)blockIdx.x * IND_WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * IND_WG) {
    table_store_private(tabs, IND_WG, threadIdx.x, eqv[i & kmask]);
    const uint4 p = comb_mul_rt(rows[i], tabs + threadIdx.x, IND_WG);
    table_store_private(tabs, IND_WG, threadIdx.x, aq[i >> log_row]);
    comb_mul_acc_rt(p, tabs + threadIdx.x, IND_WG, acc);
  }
  uint4 v = f_reduce(acc);
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = f_add(v, shfl_xor4_i(v, m));
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 s = red[0];
    for (int w = 1; w < IND_WG / 64; ++w) s = f_add(s, red[w]);
    uint32_t* o = reinterpret_cast<uint32_t*>(out);
    atomicXor(o + 0, s.x);
    atomicXor(o + 1, s.y);
    atomicXor(o + 2, s.z);
    atomicXor(o + 3, s.w);
  }
}

__global__ void zero4_kernel(uint4* p) {
  if (threadIdx.x == 0) *p = make_uint4(0u, 0u, 0u, 0u);
}

static inline uint4 h2u(flk_f128 a) {
  return make_uint4((uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32));
}

static std::vector<flk_f128> host_eq(const flk_f128* pt, size_t d) {
  std::vector<flk_f128> t(size_t(1) << d);
  t[0] = flk_f128{1, 0};
  for (size_t j = 0; j < d; ++j) {
    size_t len = size_t(1) << j;
    for (size_t i = 0; i < len; ++i) {
      flk_f128 v = t[i], hi = host_mul(v, pt[j]);
      t[i + len] = hi;
      t[i] = flk_f128{v.lo ^ hi.lo, v.hi ^ hi.hi};
    }
  }
  return t;
}

// sks_vks of eval_sk_at_vks (CF/pcs/ligerito.rs:1186-1205).
static std::vector<flk_f128> host_sks_vks(size_t log_n) {
  std::vector<flk_f128> sks(log_n + 1, flk_f128{0, 0});
  sks[0] = flk_f128{1, 0};
  if (log_n == 0) return sks;
  std::vector<flk_f128> layer;
  for (size_t i = 1; i <= log_n; ++i) layer.push_back(flk_f128{1ull << i, 0});
  size_t cur = log_n;
  for (size_t i = 0; i < log_n; ++i) {
    for (size_t j = 0; j < cur; ++j) {
      flk_f128 s = layer[j];
      flk_f128 sq = host_mul(s, s), rs = host_mul(sks[i], s);
      flk_f128 v{sq.lo ^ rs.lo, sq.hi ^ rs.hi};
      if (j == 0)
        sks[i + 1] = v;
      else
        layer[j - 1] = v;
    }
    --cur;
  }
  return sks;
}

}  // namespace flk

extern "C" int flk_induce_sumcheck_poly(size_t log_msg_cols, const flk_f128* rows, size_t n_queries,
                                        const flk_f128* v_challenges, size_t n_v, const uint64_t* queries,
                                        const flk_f128* alpha, size_t n_alpha, flk_f128* basis_out,
                                        flk_f128* sum_out, void* stream) {
  using namespace flk;
  int rc = check_enabled();
  if (rc) return rc;
  if (!basis_out || !sum_out || (n_queries && (!rows || !queries)) || (n_v && !v_challenges) ||
      (n_alpha && !alpha))
    return set_error(FLK_E_SHAPE, "induce: null pointer");
  if (log_msg_cols > 26 || n_v > 12 || n_alpha > 20 || n_queries > (size_t(1) << n_alpha) || n_queries > 65535)
    return set_error(FLK_E_SHAPE, "induce: shape out of range (cols 2^%zu, n_v %zu, queries %zu, n_alpha %zu)",
                     log_msg_cols, n_v, n_queries, n_alpha);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint32_t L = (uint32_t)log_msg_cols;
  const uint32_t nq = (uint32_t)n_queries;
  const size_t row_len = size_t(1) << n_v;
  uint4* basis = reinterpret_cast<uint4*>(basis_out);
  uint4* sum = reinterpret_cast<uint4*>(sum_out);

  hipLaunchKernelGGL(zero4_kernel, dim3(1), dim3(64), 0, st, sum);
  FLK_LAUNCH_CHECK("zero4_kernel");
  if (nq == 0) return zero_n(basis, uint64_t(1) << L, st);  // a kernel: safe inside a captured graph

  // Host constants: alpha_q, eq(v), sks_vks and their inverses (one inversion
  // for all of them: prefix products, invert the last, walk back).
  std::vector<flk_f128> aq = host_eq(alpha, n_alpha);
  aq.resize(nq);
  std::vector<flk_f128> eqv = host_eq(v_challenges, n_v);
  std::vector<flk_f128> sks = host_sks_vks(L);
  std::vector<flk_f128> consts(2 * (L + 1));
  {
    std::vector<flk_f128> pre(L + 2, flk_f128{1, 0});
    for (uint32_t i = 0; i <= L; ++i) {
      const bool nz = sks[i].lo | sks[i].hi;
      pre[i + 1] = nz ? host_mul(pre[i], sks[i]) : pre[i];
    }
    flk_f128 inv = host_inv(pre[L + 1]);  // inverse of the product of the nonzero sks
    for (uint32_t i = L + 1; i-- > 0;) {
      consts[i] = sks[i];
      if (sks[i].lo | sks[i].hi) {
        consts[L + 1 + i] = host_mul(inv, pre[i]);
        inv = host_mul(inv, sks[i]);
      } else {
        consts[L + 1 + i] = sks[i];
      }
    }
  }

  const uint32_t s_lo = L < 10 ? L : 10;
  const uint32_t s_hi = L - s_lo;
  const size_t n_consts = consts.size(), n_aq = nq, n_eqv = row_len;
  const size_t n_S = (size_t)nq * (L ? L : 1), n_lo = (size_t)nq << s_lo, n_hi = (size_t)nq << s_hi;
  const size_t total = n_consts + n_aq + n_eqv + n_S + n_lo + n_hi;
  uint4* scratch = nullptr;
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&scratch), total * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "induce: scratch allocation of %zu bytes failed", total * sizeof(uint4));
  }
  uint4* d_consts = scratch;
  uint4* d_aq = d_consts + n_consts;
  uint4* d_eqv = d_aq + n_aq;
  uint4* d_S = d_eqv + n_eqv;
  uint4* d_lo = d_S + n_S;
  uint4* d_hi = d_lo + n_lo;
  std::vector<flk_f128> up;
  up.reserve(n_consts + n_aq + n_eqv);
  up.insert(up.end(), consts.begin(), consts.end());
  up.insert(up.end(), aq.begin(), aq.end());
  up.insert(up.end(), eqv.begin(), eqv.end());
  e = upload_staged(scratch, up.data(), up.size() * sizeof(flk_f128), st);
  if (e != hipSuccess) return device_failure(e, "induce: upload constants");

  if (L > 0) {
    hipLaunchKernelGGL(induce_sks_kernel, dim3((nq + 63) / 64), dim3(64), 0, st, queries, nq, L, d_consts, d_S);
    FLK_LAUNCH_CHECK("induce_sks_kernel");
  }
  hipLaunchKernelGGL(induce_expand_kernel, dim3(nq), dim3(IND_WG), 0, st, d_S, L, 0u, s_lo,
                     static_cast<const uint4*>(nullptr), d_lo);
  FLK_LAUNCH_CHECK("induce_expand_kernel lo");
  hipLaunchKernelGGL(induce_expand_kernel, dim3(nq), dim3(IND_WG), 0, st, d_S, L, s_lo, s_hi,
                     static_cast<const uint4*>(d_aq), d_hi);
  FLK_LAUNCH_CHECK("induce_expand_kernel hi");
  // One basis entry per thread, 2^s_lo / IND_WG work-groups per jh: at the
  // largest level (2^19 entries) 2048 work-groups, eight per CU.
  const uint32_t n_lo_cols = 1u << s_lo;
  const uint32_t ny = n_lo_cols > (uint32_t)IND_WG ? n_lo_cols / IND_WG : 1u;
  hipLaunchKernelGGL(induce_basis_kernel<1>, dim3(1u << s_hi, ny), dim3(IND_WG), 0, st, d_lo, d_hi, nq, s_lo, s_hi,
                     basis);
  FLK_LAUNCH_CHECK("induce_basis_kernel");
  const uint64_t n_rows = (uint64_t)nq * row_len;
  uint64_t g = (n_rows + IND_WG - 1) / IND_WG;
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(dot_kernel, dim3((unsigned)g), dim3(IND_WG), 0, st, reinterpret_cast<const uint4*>(rows), d_eqv,
                     d_aq, (uint32_t)n_v, n_rows, sum);
  FLK_LAUNCH_CHECK("dot_kernel");
  e = flk::stream_free(scratch, st);
  if (e != hipSuccess) return device_failure(e, "induce: free scratch");
  return FLK_OK;
}
