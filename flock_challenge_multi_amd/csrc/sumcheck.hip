// Ligerito sumcheck reductions, eq tables and multilinear folds for gfx950
// (CF/pcs/ligerito.rs:1162-1175,4958-5075,5709-6058,9213-9260;
//  CF/lincheck.rs:2080-2121).
//
// Every reduction is an XOR sum, so partial sums are combined in any order:
// each lane XOR-accumulates unreduced 256-bit comb products, reduces once,
// the wave folds with cross-lane shuffles, one lane per work-group stores the
// work-group's sum as a partial, and a second launch XORs the partials
// (partials_sum_free).  The result is bit-identical to the sequential sum
// (CF/pcs.rs:96-107: schedule width cannot change wire bytes).
//
// Data x data products (f.b) use a per-lane comb table of one operand in LDS,
// laid out entry-major so a wave's 64 lanes touch 64 consecutive 16-byte slots.
// Products by a challenge (r, alpha, eq factors) use work-group-shared
// nibble-position tables (build_pos_tables, combjr_mul_add).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <mutex>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

constexpr int SC_WG = 128;

__device__ __forceinline__ uint4 shfl_xor4(uint4 v, int m) {
  return make_uint4(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m), __shfl_xor(v.w, m));
}

// XOR-reduce `v` over the work-group and store it at parts[k * gridDim.x +
// blockIdx.x] (the work-group partials that partials_sum_free adds up).
__device__ __forceinline__ void wg_xor_reduce_store(uint4 v, uint4* red_lds, uint4* parts, uint32_t k) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = f_add(v, shfl_xor4(v, m));
  const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  if (lane == 0) red_lds[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 acc = red_lds[0];
    for (uint32_t w = 1; w < (blockDim.x + 63u) / 64u; ++w) acc = f_add(acc, red_lds[w]);
    parts[(size_t)k * gridDim.x + blockIdx.x] = acc;
  }
}

// ---------------------------------------------------------------------------
// round_msg_lsb / round_msg_and_eval_lsb
// ---------------------------------------------------------------------------
// out: work-group partials (2 or 3 outputs x gridDim.x), summed afterwards.
template <bool EVAL>
__global__ void __launch_bounds__(SC_WG) round_msg_kernel(const uint4* __restrict__ f, const uint4* __restrict__ b,
                                                         uint64_t n_pairs, ReduceTail rt) {
  __shared__ uint4 tabs[8 * SC_WG];
  __shared__ uint4 red[3][SC_WG / 64];
  const uint32_t tid = threadIdx.x;
  U256 a0, a2, a1;
  u256_zero(a0);
  u256_zero(a2);
  u256_zero(a1);
  uint4* my = tabs + tid;
  table3_init(tabs, tid);
  for (uint64_t j = (uint64_t)blockIdx.x * SC_WG + tid; j < n_pairs; j += (uint64_t)gridDim.x * SC_WG) {
    const uint4 f0 = f[2 * j], f1 = f[2 * j + 1], b0 = b[2 * j], b1 = b[2 * j + 1];
    table3_store_private(tabs, SC_WG, tid, f0);
    comb3_mul_acc_rt(b0, my, SC_WG, a0);
    table3_store_private(tabs, SC_WG, tid, f_add(f0, f1));
    comb3_mul_acc_rt(f_add(b0, b1), my, SC_WG, a2);
    if (EVAL) {
      table3_store_private(tabs, SC_WG, tid, f1);
      comb3_mul_acc_rt(b1, my, SC_WG, a1);
    }
  }
  const uint4 u0 = f_reduce(a0), u2 = f_reduce(a2);
  wg_xor_reduce_store(u0, red[0], rt.parts, 0);
  wg_xor_reduce_store(u2, red[1], rt.parts, 1);
  if (EVAL) wg_xor_reduce_store(f_add(u0, f_reduce(a1)), red[2], rt.parts, 2);
}

// Work-groups of the product-sum kernels (sums go through partials, so the
// count costs no atomics); FLOCK_SC_GRID overrides it for sweeps.
static unsigned reduce_grid(uint64_t items) {
  static const uint64_t cap = [] {
    const char* v = getenv("FLOCK_SC_GRID");
    const long k = v ? atol(v) : 8192;
    return (uint64_t)(k < 1 ? 1 : (k > 65535 ? 65535 : k));
  }();
  uint64_t g = (items + SC_WG - 1) / SC_WG;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return (unsigned)g;
}

// ---------------------------------------------------------------------------
// fold_and_msg_lsb: one lane owns one output pair (4 inputs of f and b).
// ---------------------------------------------------------------------------
// out: work-group partials (2 outputs x gridDim.x), summed afterwards.
__global__ void __launch_bounds__(SC_WG) fold_and_msg_kernel(const uint4* __restrict__ f, const uint4* __restrict__ b,
                                                            uint64_t n_quads, uint4 r, uint4* __restrict__ f_out,
                                                            uint4* __restrict__ b_out, ReduceTail rt) {
  __shared__ uint4 tabs[8 * SC_WG];
  __shared__ uint4 rtab[128];  // nibble-position tables of r
  __shared__ uint4 rnib[32];
  __shared__ uint4 red[2][SC_WG / 64];
  const uint32_t tid = threadIdx.x;
  build_pos_tables(rtab, rnib, 1, [&](uint32_t) { return r; });
  U256 a0, a2;
  u256_zero(a0);
  u256_zero(a2);
  uint4* my = tabs + tid;
  table3_init(tabs, tid);
  for (uint64_t q = (uint64_t)blockIdx.x * SC_WG + tid; q < n_quads; q += (uint64_t)gridDim.x * SC_WG) {
    const uint4* fp = f + 4 * q;
    const uint4* bp = b + 4 * q;
    const uint4 fa = fp[0], fb = fp[1], fc = fp[2], fd = fp[3];
    const uint4 ba = bp[0], bb = bp[1], bc = bp[2], bd = bp[3];
    const uint4 nf0 = combjr_mul_add(f_add(fa, fb), fa, rtab);
    const uint4 nf1 = combjr_mul_add(f_add(fc, fd), fc, rtab);
    const uint4 nb0 = combjr_mul_add(f_add(ba, bb), ba, rtab);
    const uint4 nb1 = combjr_mul_add(f_add(bc, bd), bc, rtab);
    f_out[2 * q] = nf0;
    f_out[2 * q + 1] = nf1;
    b_out[2 * q] = nb0;
    b_out[2 * q + 1] = nb1;
    table3_store_private(tabs, SC_WG, tid, nf0);
    comb3_mul_acc_rt(nb0, my, SC_WG, a0);
    table3_store_private(tabs, SC_WG, tid, f_add(nf0, nf1));
    comb3_mul_acc_rt(f_add(nb0, nb1), my, SC_WG, a2);
  }
  wg_xor_reduce_store(f_reduce(a0), red[0], rt.parts, 0);
  wg_xor_reduce_store(f_reduce(a2), red[1], rt.parts, 1);
}

// ---------------------------------------------------------------------------
// glue: b += alpha * b_new
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) glue_kernel(uint4* __restrict__ b, const uint4* __restrict__ bn, uint64_t n,
                                                  uint4 alpha) {
  __shared__ uint4 atab[128];  // nibble-position tables of alpha
  __shared__ uint4 anib[32];
  build_pos_tables(atab, anib, 1, [&](uint32_t) { return alpha; });
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256)
    b[i] = combjr_mul_add(bn[i], b[i], atab);
}

// ---------------------------------------------------------------------------
// eq table: out[i] = prod_j (bit_j(i) ? r_j : 1 + r_j)  (LSB first)
// ---------------------------------------------------------------------------
struct EqFactors {
  uint4 e[16];  // e[k] = prod over the pass's t coordinates, selected by the bits of k
};

// First 2^s entries (s <= 8) by one work-group: entry i = lo[i & 15] * hi[i >> 4]
// (hi only when s > 4), where lo and hi hold the products over the first
// min(s, 4) coordinates and over coordinates 4 .. s-1 for every bit pattern:
// one product per entry, on the lane's private comb table.
__global__ void __launch_bounds__(256) eq_seed_kernel(uint4* __restrict__ out, EqFactors lo, EqFactors hi,
                                                      uint32_t s) {
  __shared__ uint4 tabs[16 * 256];
  const uint32_t i = threadIdx.x;
  if (i >= (1u << s)) return;  // no barrier below: each lane reads only its own table
  uint4 v = lo.e[i & 15u];
  if (s > 4) {
    table_store_private(tabs, blockDim.x, i, hi.e[i >> 4]);
    v = comb_mul_rt(v, tabs + i, blockDim.x);
  }
  out[i] = v;
}

// out[i + k*len] = out[i] * E[k] for k < 2^t, i < len (in place; reads [0,len)).
// The 2^t factors get nibble-position tables (2 KiB each); each lane computes
// its value's 32 table offsets once and runs the 2^t shift-free products.
__global__ void __launch_bounds__(256) eq_expand_kernel(uint4* __restrict__ out, uint64_t len, uint32_t t, EqFactors E) {
  __shared__ uint4 nib[16 * 32];    // E[c] * x^s
  __shared__ uint4 tabs[16 * 128];  // [c][j][k] = k * E[c] * x^(4j)
  const uint32_t nk = 1u << t;
  for (uint32_t e = threadIdx.x; e < nk * 32u; e += 256) nib[e] = f_mul_xs(E.e[e >> 5], e & 31u);
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < nk * 128u; e += 256) {
    const uint32_t c = e >> 7, j = (e >> 4) & 7u, k = e & 15u;
    uint4 v = f_zero();
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if ((k >> b) & 1u) v = f_add(v, nib[c * 32 + 4 * j + b]);
    tabs[e] = v;
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < len; i += (uint64_t)gridDim.x * 256) {
    uint32_t ad[32];
    pos_offsets(out[i], ad);
    for (uint32_t k = 0; k < nk; ++k) out[i + k * len] = combjr_mul_pre(ad, tabs + 128 * k);
  }
}

// ---------------------------------------------------------------------------
// partial_eval_lsb: T folds per pass, out[i] = e0 + r (e0 + e1) each fold.
// ---------------------------------------------------------------------------
template <int T>
__global__ void __launch_bounds__(256) fold_pass_kernel(const uint4* __restrict__ in, uint4* __restrict__ out,
                                                       uint64_t n_out, EqFactors rs) {
  __shared__ uint4 tabs[128 * T];  // nibble-position tables of the T challenges
  __shared__ uint4 nib[32 * T];
  build_pos_tables(tabs, nib, T, [&](uint32_t c) { return rs.e[c]; });
  const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n_out) return;
  uint4 x[1 << T];
#pragma unroll
  for (int j = 0; j < (1 << T); ++j) x[j] = in[(i << T) + j];
#pragma unroll
  for (int t = 0; t < T; ++t) {
#pragma unroll
    for (int j = 0; j < (1 << (T - 1 - t)); ++j) {
      const uint4 e0 = x[2 * j], e1 = x[2 * j + 1];
      x[j] = combjr_mul_add(f_add(e0, e1), e0, tabs + 128 * t);
    }
  }
  out[i] = x[0];
}

// Scratch of the allocation-free round message (flk_sc_round_msg_noalloc): the
// work-group partials of at most NOALLOC_GRID work-groups for 2 outputs
// (128 KiB per device, held until the process ends).  Host code never uses the
// symbol's name as an address: noalloc_scratch() resolves each device's copy
// once with hipGetSymbolAddress and every launch receives that pointer.
constexpr unsigned NOALLOC_GRID = 4096;
static __device__ __align__(256) uint4 scratch[2 * NOALLOC_GRID];

static uint4* noalloc_scratch(int* status) {
  static std::mutex mu;
  static uint4* per_dev[64] = {};
  *status = FLK_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = device_failure(e, "hipGetDevice");
    return nullptr;
  }
  if (dev < 0 || dev >= 64) {
    *status = set_error(FLK_E_SHAPE, "round_msg_noalloc: device index %d out of range", dev);
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(mu);
  if (!per_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      *status = device_failure(e, "hipGetSymbolAddress(scratch)");
      return nullptr;
    }
    per_dev[dev] = static_cast<uint4*>(p);
  }
  return per_dev[dev];
}

static inline uint4 to_u4(flk_f128 a) {
  return make_uint4((uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32));
}

}  // namespace flk

using flk::SC_WG;

extern "C" {

int flk_sc_round_msg(const flk_f128* f, const flk_f128* b, size_t n, flk_f128* out2, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!f || !b || !out2) return flk::set_error(FLK_E_SHAPE, "round_msg: null pointer");
  if (!flk::is_pow2(n) || n < 2) return flk::set_error(FLK_E_SHAPE, "round_msg: n=%zu must be a power of two >= 2", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out2);
  const uint64_t pairs = n / 2;
  const unsigned grid = flk::reduce_grid(pairs);
  flk::ReduceTail rt;
  if ((rc = flk::reduce_begin(grid, 2, o, st, &rt))) return rc;
  hipLaunchKernelGGL(flk::round_msg_kernel<false>, dim3(grid), dim3(SC_WG), 0, st,
                     reinterpret_cast<const uint4*>(f), reinterpret_cast<const uint4*>(b), pairs, rt);
  FLK_LAUNCH_CHECK("round_msg_kernel");
  return flk::reduce_end(rt, grid, 2, st);
}

int flk_sc_round_msg_noalloc(const flk_f128* f, const flk_f128* b, size_t n, flk_f128* out2, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!f || !b || !out2) return flk::set_error(FLK_E_SHAPE, "round_msg_noalloc: null pointer");
  if (!flk::is_pow2(n) || n < 2 || n > (size_t(1) << 40))
    return flk::set_error(FLK_E_SHAPE, "round_msg_noalloc: n=%zu must be a power of two in [2, 2^40]", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out2);
  const uint64_t pairs = n / 2;
  uint64_t g = (pairs + SC_WG - 1) / SC_WG;
  const unsigned grid = (unsigned)(g > flk::NOALLOC_GRID ? flk::NOALLOC_GRID : g);
  uint4* parts = o;  // one work-group writes out directly
  if (grid > 1) {
    if (grid > flk::NOALLOC_GRID)
      return flk::set_error(FLK_E_SHAPE, "round_msg_noalloc: %u partials exceed the %u-entry scratch", grid,
                            flk::NOALLOC_GRID);
    parts = flk::noalloc_scratch(&rc);
    if (!parts) return rc;
  }
  const flk::ReduceTail rt{parts, o, nullptr, nullptr, 0u};
  hipLaunchKernelGGL(flk::round_msg_kernel<false>, dim3(grid), dim3(SC_WG), 0, st,
                     reinterpret_cast<const uint4*>(f), reinterpret_cast<const uint4*>(b), pairs, rt);
  FLK_LAUNCH_CHECK("round_msg_kernel(noalloc)");
  if (grid > 1) {
    // every partial [k * grid, k * grid + grid) was just stored by its work-group,
    // so nothing in the scratch has to start at zero
    if ((rc = flk::partials_sum(parts, grid, 2, o, st))) return rc;
  }
  return FLK_OK;
}

int flk_sc_round_msg_and_eval(const flk_f128* f, const flk_f128* b, size_t n, flk_f128* out3, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!f || !b || !out3) return flk::set_error(FLK_E_SHAPE, "round_msg_and_eval: null pointer");
  if (!flk::is_pow2(n) || n < 2)
    return flk::set_error(FLK_E_SHAPE, "round_msg_and_eval: n=%zu must be a power of two >= 2", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out3);
  const uint64_t pairs = n / 2;
  const unsigned grid = flk::reduce_grid(pairs);
  flk::ReduceTail rt;
  if ((rc = flk::reduce_begin(grid, 3, o, st, &rt))) return rc;
  hipLaunchKernelGGL(flk::round_msg_kernel<true>, dim3(grid), dim3(SC_WG), 0, st,
                     reinterpret_cast<const uint4*>(f), reinterpret_cast<const uint4*>(b), pairs, rt);
  FLK_LAUNCH_CHECK("round_msg_kernel<eval>");
  return flk::reduce_end(rt, grid, 3, st);
}

int flk_sc_fold_and_msg(const flk_f128* f, const flk_f128* b, size_t n, flk_f128 r, flk_f128* f_out, flk_f128* b_out,
                        flk_f128* out2, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!f || !b || !f_out || !b_out || !out2) return flk::set_error(FLK_E_SHAPE, "fold_and_msg: null pointer");
  if (!flk::is_pow2(n) || n < 4) return flk::set_error(FLK_E_SHAPE, "fold_and_msg: n=%zu must be a power of two >= 4", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out2);
  const uint64_t quads = n / 4;
  const unsigned grid = flk::reduce_grid(quads);
  flk::ReduceTail rt;
  if ((rc = flk::reduce_begin(grid, 2, o, st, &rt))) return rc;
  hipLaunchKernelGGL(flk::fold_and_msg_kernel, dim3(grid), dim3(SC_WG), 0, st,
                     reinterpret_cast<const uint4*>(f), reinterpret_cast<const uint4*>(b), quads, flk::to_u4(r),
                     reinterpret_cast<uint4*>(f_out), reinterpret_cast<uint4*>(b_out), rt);
  FLK_LAUNCH_CHECK("fold_and_msg_kernel");
  return flk::reduce_end(rt, grid, 2, st);
}

int flk_sc_glue(flk_f128* b, const flk_f128* b_new, size_t n, flk_f128 alpha, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!b || !b_new) return flk::set_error(FLK_E_SHAPE, "glue: null pointer");
  if (n == 0) return FLK_OK;
  uint64_t g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  hipLaunchKernelGGL(flk::glue_kernel, dim3((unsigned)g), dim3(256), 0, static_cast<hipStream_t>(stream),
                     reinterpret_cast<uint4*>(b), reinterpret_cast<const uint4*>(b_new), (uint64_t)n, flk::to_u4(alpha));
  FLK_LAUNCH_CHECK("glue_kernel");
  return FLK_OK;
}

int flk_build_eq(const flk_f128* point, size_t d, flk_f128* out, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!out || (d && !point)) return flk::set_error(FLK_E_SHAPE, "build_eq: null pointer");
  if (d > 34) return flk::set_error(FLK_E_SHAPE, "build_eq: d=%zu too large", d);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out);
  const size_t s = d < 8 ? d : 8;
  // Products over point[first .. first + t) for every bit pattern k < 2^t.
  auto combos = [&](size_t first, size_t t) {
    flk::EqFactors E{};
    for (size_t k = 0; k < (size_t(1) << t); ++k) {
      flk_f128 acc{1, 0};
      for (size_t bit = 0; bit < t; ++bit) {
        const flk_f128 r = point[first + bit];
        acc = flk::host_mul(acc, ((k >> bit) & 1) ? r : flk_f128{r.lo ^ 1ull, r.hi});
      }
      E.e[k] = flk::to_u4(acc);
    }
    return E;
  };
  const size_t t_lo = s < 4 ? s : 4;
  const flk::EqFactors lo = combos(0, t_lo);
  const flk::EqFactors hi = s > 4 ? combos(4, s - 4) : flk::EqFactors{};
  hipLaunchKernelGGL(flk::eq_seed_kernel, dim3(1), dim3(256), 0, st, o, lo, hi, (uint32_t)s);
  FLK_LAUNCH_CHECK("eq_seed_kernel");
  if (d == 0) return FLK_OK;
  size_t done = s;
  while (done < d) {
    const size_t t = (d - done) < 4 ? (d - done) : 4;
    const flk::EqFactors E = combos(done, t);
    const uint64_t len = uint64_t(1) << done;
    // Each work-group builds 2^t tables of 2 KiB, so it strides over several
    // inputs: at most 1024 work-groups.
    const uint64_t blocks = (len + 255) / 256 < 1024 ? (len + 255) / 256 : 1024;
    hipLaunchKernelGGL(flk::eq_expand_kernel, dim3((unsigned)blocks), dim3(256), 0, st, o, len, (uint32_t)t, E);
    FLK_LAUNCH_CHECK("eq_expand_kernel");
    done += t;
  }
  return FLK_OK;
}

int flk_partial_eval_lsb(const flk_f128* evals, size_t n, const flk_f128* rs, size_t k, flk_f128* out, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!evals || !out || (k && !rs)) return flk::set_error(FLK_E_SHAPE, "partial_eval: null pointer");
  if (!flk::is_pow2(n) || (size_t(1) << k) > n)
    return flk::set_error(FLK_E_SHAPE, "partial_eval: n=%zu k=%zu", n, k);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (k == 0) {
    hipError_t e = hipMemcpyAsync(out, evals, n * sizeof(flk_f128), hipMemcpyDeviceToDevice, st);
    if (e != hipSuccess) return flk::device_failure(e, "partial_eval copy");
    return FLK_OK;
  }
  const uint4* src = reinterpret_cast<const uint4*>(evals);
  const int n_passes = (int)((k + 3) / 4);
  // Passes never fold in place: intermediates ping-pong between two scratch
  // regions (n/16 and n/256 elements) and the last pass writes `out`.
  uint4* tmp = nullptr;
  if (n_passes > 1) {
    hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&tmp), ((n >> 4) + (n >> 8) + 1) * sizeof(uint4), st);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return flk::set_error(FLK_E_NOMEM, "partial_eval: scratch allocation failed");
    }
  }
  uint4* scratch[2] = {tmp, tmp ? tmp + (n >> 4) : nullptr};
  size_t cur = n, done = 0;
  for (int pass = 0; pass < n_passes; ++pass) {
    const size_t t = (k - done) < 4 ? (k - done) : 4;
    flk::EqFactors R{};
    for (size_t j = 0; j < t; ++j) R.e[j] = flk::to_u4(rs[done + j]);
    const uint64_t n_out = cur >> t;
    uint4* dst = (pass == n_passes - 1) ? reinterpret_cast<uint4*>(out) : scratch[pass & 1];
    const unsigned g = (unsigned)((n_out + 255) / 256);
    switch (t) {
      case 1: hipLaunchKernelGGL(flk::fold_pass_kernel<1>, dim3(g), dim3(256), 0, st, src, dst, n_out, R); break;
      case 2: hipLaunchKernelGGL(flk::fold_pass_kernel<2>, dim3(g), dim3(256), 0, st, src, dst, n_out, R); break;
      case 3: hipLaunchKernelGGL(flk::fold_pass_kernel<3>, dim3(g), dim3(256), 0, st, src, dst, n_out, R); break;
      default: hipLaunchKernelGGL(flk::fold_pass_kernel<4>, dim3(g), dim3(256), 0, st, src, dst, n_out, R); break;
    }
    FLK_LAUNCH_CHECK("fold_pass_kernel");
    src = dst;
    cur = n_out;
    done += t;
  }
  if (tmp) {
    hipError_t e = flk::stream_free(tmp, st);
    if (e != hipSuccess) return flk::device_failure(e, "partial_eval free");
  }
  return FLK_OK;
}

}  // extern "C"
