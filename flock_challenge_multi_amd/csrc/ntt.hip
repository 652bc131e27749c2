// Interleaved additive NTT / Reed-Solomon encode over GF(2^128) for gfx950.
//
// Semantics: AdditiveNttF128::forward_transform_interleaved_from_layer and
// rs_encode_interleaved (CF/ntt/additive_ntt_f128.rs:2728-2830; scalar form
// :4766-4795).  Layout data[pos * num_ntts + lane]; layer l pairs positions
// 2^(log_d-l-1) apart inside blocks of 2^(log_d-l); u += v*twiddle(l,b); v += u.
//
// Schedule.  The layers are processed in passes of K consecutive layers.  A
// thread owns one lane of one "set": the 2^K positions that the K layers of the
// pass connect (they differ only in the K position bits the pass consumes), so a
// pass reads and writes every element once and all K layers run in registers.
// Consecutive threads are consecutive lanes of the same positions, so every
// access is a contiguous 16-byte-per-lane stream.
//
// Twiddle products are combs (gf128.hpp).  In every pass but the deepest, all
// sets of a work-group share their twiddles and the work-group builds their
// nibble-position tables once (combjr_mul_add); each thread runs two sets on
// them.  In the deepest passes a set's twiddles are its own: with 64 lanes or
// more one wave (= one work-group) owns a set and builds the nibble-position
// tables of each layer's twiddles (ntt_pass_pos_kernel); with fewer lanes the
// G = num_ntts lanes of a set build nibble-pair tables (comb4p_mul_add).  The
// inverse transform runs the same passes in reverse order with the inverse
// butterfly (v += u; u += v*twiddle).
//
// The RS encode reads the message directly in its first pass: after the
// log_inv_rate copy-only layers of the zero-padded transform, position p holds
// message position p mod 2^(log_d - log_inv_rate) (replicate_message_fill,
// :7455-7476), so the replicated codeword is never materialised.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

constexpr int NTT_WG = 256;
// Occupancy target of the bitop3 comb kernels: 4 waves per SIMD (<= 128 VGPRs);
// measured faster than 2 waves with the comb unrolled or 5 waves with spills.
constexpr int NTT_WAVES = 4;

// PAIR: 32-entry nibble-pair tables and comb4p_mul_add.  INV: the inverse
// transform's pass (layers deepest first, butterfly v += u; u += v * t).
template <int K, bool PAIR, bool INV = false>
__global__ void __launch_bounds__(NTT_WG) __attribute__((amdgpu_waves_per_eu(PAIR && K <= 3 ? NTT_WAVES : 1))) ntt_pass_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                         const uint4* __restrict__ tw, uint32_t log_d, uint32_t L,
                                                         uint32_t log_lanes, uint64_t src_pos_mask,
                                                         uint64_t n_threads, uint32_t log_group) {
  extern __shared__ uint4 lds[];
  const uint32_t G = 1u << log_group;            // lanes sharing one set's twiddles (<= 64)
  const uint32_t groups_per_wave = 64u >> log_group;
  const uint32_t tid = threadIdx.x;
  const uint32_t wave = tid >> 6, lane_w = tid & 63u;
  const uint32_t grp = lane_w >> log_group, lane_g = lane_w & (G - 1u);
  constexpr int TAB = PAIR ? 32 : 16;  // entries per twiddle
  uint4* my_tabs = lds + (size_t)(wave * groups_per_wave + grp) * ((size_t)TAB << (K - 1));

  const uint64_t g = (uint64_t)blockIdx.x * NTT_WG + tid;
  const bool active = g < n_threads;
  const uint64_t c = g & ((1ull << log_lanes) - 1ull);
  const uint64_t s = g >> log_lanes;
  const uint32_t lo_bits = log_d - L - K;
  const uint64_t lo = s & ((1ull << lo_bits) - 1ull);
  const uint64_t prefix = s >> lo_bits;

  uint4 x[1 << K];
#pragma unroll
  for (int j = 0; j < (1 << K); ++j) {
    const uint64_t p = (prefix << (log_d - L)) | ((uint64_t)j << lo_bits) | lo;
    x[j] = active ? src[((p & src_pos_mask) << log_lanes) | c] : f_zero();
  }

#pragma unroll
  for (int ii = 0; ii < K; ++ii) {
    const int i = INV ? K - 1 - ii : ii;
    // Tables of the 2^i twiddles of layer L+i met by this set.
    const uint32_t n_entries = (uint32_t)TAB << i;
    const uint64_t tw_base = ((1ull << (L + i)) - 1ull) + (prefix << i);
    for (uint32_t e = lane_g; e < n_entries; e += G) {
      const uint32_t sb = PAIR ? e >> 5 : e >> 4, k = e & 15u;
      uint4 t = active ? tw[tw_base + sb] : f_zero();
      if (PAIR && ((e >> 4) & 1u)) t = f_mul_x4(t);
      my_tabs[e] = comb_entry(t, k);
    }
    __syncthreads();
    constexpr int dummy = 0;
    (void)dummy;
#pragma unroll
    for (int sb = 0; sb < (1 << i); ++sb) {
      const uint4* tab = my_tabs + sb * TAB;
#pragma unroll
      for (int jj = 0; jj < (1 << (K - 1 - i)); ++jj) {
        const int j0 = (sb << (K - i)) + jj;
        const int j1 = j0 + (1 << (K - 1 - i));
        if constexpr (INV) x[j1] = f_add(x[j1], x[j0]);
        if constexpr (PAIR) {
          x[j0] = comb4p_mul_add(x[j1], x[j0], tab);
        } else {
          const uint4 prod = comb_mul<1>(x[j1], tab);
          x[j0] = f_add(x[j0], prod);
        }
        if constexpr (!INV) x[j1] = f_add(x[j1], x[j0]);
      }
    }
    __syncthreads();
  }

  if (active) {
#pragma unroll
    for (int j = 0; j < (1 << K); ++j) {
      const uint64_t p = (prefix << (log_d - L)) | ((uint64_t)j << lo_bits) | lo;
      dst[(p << log_lanes) | c] = x[j];
    }
  }
}

// Deepest passes with one set per wave (num_ntts >= 64), one wave per
// work-group: before each layer the wave builds the nibble-position tables
// (k * t * x^(4j), j = 0..7; 2 KiB per twiddle) of the 2^i twiddles that layer
// meets, so every product is the shift-free combjr_mul_add.  The tables sit at
// LDS address 0 of the work-group, so every lookup address is the nibble alone
// plus an immediate offset.  Table build: lane item (sb, j, half) forms
// b_i = t_sb * x^(4j+i), i = 0..3, and writes entries half*8 .. half*8+7 as
// XORs of them (Gray order, one XOR per entry).  Position j's table starts at
// entry 17j (one pad entry), so the 16 lanes of each write cover all 16 LDS
// bank groups; with 16-entry strides every write hit two of them (0.21 of the
// CU cycles in bank conflicts).  The twiddles of all K layers are loaded with
// the data, before the first table build.
constexpr int NTT_POS_WG = 64;
template <int K, bool INV = false>
__global__ void __launch_bounds__(NTT_POS_WG) __attribute__((amdgpu_waves_per_eu(NTT_WAVES)))
ntt_pass_pos_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, const uint4* __restrict__ tw,
                    uint32_t log_d, uint32_t L, uint32_t log_lanes, uint64_t src_pos_mask, uint64_t n_threads) {
  constexpr int JS = 17;         // entries per nibble-position table (16 + 1 pad)
  constexpr int TAB = 8 * JS;    // entries per twiddle
  __shared__ uint4 tabs[TAB << (K - 1)];
  const uint32_t lane = threadIdx.x;

  const uint64_t g = (uint64_t)blockIdx.x * NTT_POS_WG + lane;
  const bool active = g < n_threads;
  const uint64_t c = g & ((1ull << log_lanes) - 1ull);
  const uint64_t s = g >> log_lanes;
  const uint32_t lo_bits = log_d - L - K;
  const uint64_t lo = s & ((1ull << lo_bits) - 1ull);
  const uint64_t prefix = s >> lo_bits;

  uint4 x[1 << K];
#pragma unroll
  for (int j = 0; j < (1 << K); ++j) {
    const uint64_t p = (prefix << (log_d - L)) | ((uint64_t)j << lo_bits) | lo;
    x[j] = active ? src[((p & src_pos_mask) << log_lanes) | c] : f_zero();
  }
  // Item of lane `lane` in layer i: twiddle sb = lane >> 4 (lanes < 16 << i).
  uint4 t_item[K];
#pragma unroll
  for (int i = 0; i < K; ++i) {
    const uint32_t sb = lane >> 4;
    t_item[i] = (active && lane < (16u << i)) ? tw[((1ull << (L + i)) - 1ull) + (prefix << i) + sb] : f_zero();
  }

#pragma unroll
  for (int ii = 0; ii < K; ++ii) {
    const int i = INV ? K - 1 - ii : ii;
    if (lane < (16u << i)) {
      const uint32_t sb = lane >> 4, j = (lane >> 1) & 7u, half = lane & 1u;
      const uint4 b0 = f_mul_xs(t_item[i], 4u * j);
      const uint4 b1 = f_mul_x(b0), b2 = f_mul_x(b1);
      uint4 e = half ? f_mul_x(b2) : f_zero();  // entry half*8 + 0
      uint4* out = tabs + sb * TAB + j * JS + half * 8u;
      // Gray order over the low three index bits: 0,1,3,2,6,7,5,4.
      out[0] = e;
      e = f_add(e, b0); out[1] = e;
      e = f_add(e, b1); out[3] = e;
      e = f_add(e, b0); out[2] = e;
      e = f_add(e, b2); out[6] = e;
      e = f_add(e, b0); out[7] = e;
      e = f_add(e, b1); out[5] = e;
      e = f_add(e, b0); out[4] = e;
    }
    __syncthreads();
#pragma unroll
    for (int sb = 0; sb < (1 << i); ++sb) {
      const uint4* tab = tabs + sb * TAB;
#pragma unroll
      for (int jj = 0; jj < (1 << (K - 1 - i)); ++jj) {
        const int j0 = (sb << (K - i)) + jj;
        const int j1 = j0 + (1 << (K - 1 - i));
        if constexpr (INV) x[j1] = f_add(x[j1], x[j0]);
        x[j0] = combjr_mul_add<JS>(x[j1], x[j0], tab);
        if constexpr (!INV) x[j1] = f_add(x[j1], x[j0]);
      }
    }
    __syncthreads();
  }

  if (active) {
#pragma unroll
    for (int j = 0; j < (1 << K); ++j) {
      const uint64_t p = (prefix << (log_d - L)) | ((uint64_t)j << lo_bits) | lo;
      dst[(p << log_lanes) | c] = x[j];
    }
  }
}

// Passes whose work-group covers sets that all share the block prefix (every
// pass except the deepest ones): the 2^K - 1 twiddles of the pass are common to
// the whole work-group, so their comb tables are built once per work-group at
// kernel start.
//   COMB 0: 256-entry byte tables (4 KiB per twiddle), two-input comb8_mul;
//   COMB 1: the same tables, comb8r_mul_add (three-input XORs, rolled);
//   COMB 2: eight 16-entry nibble-position tables (2 KiB per twiddle),
//           combjr_mul_add: twice the lookups, none of them bank-conflicted.
// Byte tables: the nibble tables of t and t*x^4 first, then entry
// b = T_lo[b & 15] ^ T_hi[b >> 4].  Nibble-position tables: the 32 values
// t*x^s first, then entry (j, k) = XOR of t*x^(4j+i) over the bits i of k.
// INV: the inverse transform's pass (layers deepest first, v += u; u += v * t).
// SETS: sets per thread; thread t of work-group w runs the sets of flat thread
// index w * 256 * SETS + r * 256 + t for r = 0 .. SETS-1 in turn (all of them
// share the work-group's prefix, checked at launch).
template <int K, int COMB, bool INV = false, int SETS = 1>
__global__ void __launch_bounds__(NTT_WG) __attribute__((amdgpu_waves_per_eu(COMB != 0 && K <= 3 ? NTT_WAVES : 1)))
ntt_pass_shared_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst, const uint4* __restrict__ tw,
                       uint32_t log_d, uint32_t L, uint32_t log_lanes, uint64_t src_pos_mask, uint64_t n_threads) {
  constexpr int NTW = (1 << K) - 1;
  constexpr int TAB = COMB == 2 ? 128 : 256;  // entries per twiddle
  __shared__ uint4 tabs[NTW * TAB];
  __shared__ uint4 nib[NTW * 32];
  const uint32_t tid = threadIdx.x;
  const uint32_t lo_bits = log_d - L - K;
  // Every set of this work-group (SETS per thread, run one after the other on
  // the same tables) has the same prefix; take it from the first thread's set.
  const uint64_t g0 = (uint64_t)blockIdx.x * (NTT_WG * SETS);
  const uint64_t prefix = (g0 >> log_lanes) >> lo_bits;
  const uint64_t stride = 1ull << (lo_bits + log_lanes);
  uint64_t g = g0 + tid;
  bool active = g < n_threads;
  // element (p_j, c) with p_j = prefix | j | lo: base + j * stride
  uint64_t base = ((((prefix << (log_d - L)) | ((g >> log_lanes) & ((1ull << lo_bits) - 1ull)))) << log_lanes) |
                  (g & ((1ull << log_lanes) - 1ull));
  uint64_t sbase = (((base >> log_lanes) & src_pos_mask) << log_lanes) | (g & ((1ull << log_lanes) - 1ull));

  // Loads first so their latency overlaps the table build.
  uint4 x[1 << K];
#pragma unroll
  for (int j = 0; j < (1 << K); ++j)
    x[j] = active ? src[sbase + ((((uint64_t)j << lo_bits) & src_pos_mask) << log_lanes)] : f_zero();

  // Twiddle q (0 .. 2^K - 2) is layer L+i, sub-block sb with q = 2^i - 1 + sb.
  if constexpr (COMB == 2) {
    for (uint32_t e = tid; e < (uint32_t)NTW * 32u; e += NTT_WG) {
      const uint32_t q = e >> 5, sh = e & 31u;
      const uint32_t i = 31u - __builtin_clz(q + 1u);
      const uint32_t sb = q + 1u - (1u << i);
      nib[e] = f_mul_xs(tw[((1ull << (L + i)) - 1ull) + (prefix << i) + sb], sh);
    }
    __syncthreads();
    for (uint32_t e = tid; e < (uint32_t)NTW * 128u; e += NTT_WG) {
      const uint32_t q = e >> 7, j = (e >> 4) & 7u, k = e & 15u;
      uint4 v = f_zero();
#pragma unroll
      for (int b = 0; b < 4; ++b)
        if ((k >> b) & 1u) v = f_add(v, nib[q * 32 + 4 * j + b]);
      tabs[e] = v;
    }
  } else {
    for (uint32_t e = tid; e < (uint32_t)NTW * 32u; e += NTT_WG) {
      const uint32_t q = e >> 5, half = (e >> 4) & 1u, k = e & 15u;
      const uint32_t i = 31u - __builtin_clz(q + 1u);
      const uint32_t sb = q + 1u - (1u << i);
      uint4 t = tw[((1ull << (L + i)) - 1ull) + (prefix << i) + sb];
      if (half) t = f_mul_x4(t);
      nib[e] = comb_entry(t, k);
    }
    __syncthreads();
    for (uint32_t e = tid; e < (uint32_t)NTW * 256u; e += NTT_WG) {
      const uint32_t q = e >> 8, b = e & 255u;
      tabs[e] = f_add(nib[q * 32 + (b & 15u)], nib[q * 32 + 16 + (b >> 4)]);
    }
  }
  __syncthreads();

#pragma unroll 1
  for (int rep = 0; rep < SETS; ++rep) {
  if (rep > 0) {
    g += NTT_WG;
    active = g < n_threads;
    base = ((((prefix << (log_d - L)) | ((g >> log_lanes) & ((1ull << lo_bits) - 1ull)))) << log_lanes) |
           (g & ((1ull << log_lanes) - 1ull));
    sbase = (((base >> log_lanes) & src_pos_mask) << log_lanes) | (g & ((1ull << log_lanes) - 1ull));
#pragma unroll
    for (int j = 0; j < (1 << K); ++j)
      x[j] = active ? src[sbase + ((((uint64_t)j << lo_bits) & src_pos_mask) << log_lanes)] : f_zero();
  }
#pragma unroll
  for (int ii = 0; ii < K; ++ii) {
    const int i = INV ? K - 1 - ii : ii;
#pragma unroll
    for (int sb = 0; sb < (1 << i); ++sb) {
      const uint4* tab = tabs + ((1 << i) - 1 + sb) * TAB;
#pragma unroll
      for (int jj = 0; jj < (1 << (K - 1 - i)); ++jj) {
        const int j0 = (sb << (K - i)) + jj;
        const int j1 = j0 + (1 << (K - 1 - i));
        if constexpr (INV) x[j1] = f_add(x[j1], x[j0]);
        if constexpr (COMB == 2)
          x[j0] = combjr_mul_add(x[j1], x[j0], tab);
        else if constexpr (COMB == 1)
          x[j0] = comb8r_mul_add(x[j1], x[j0], tab);
        else
          x[j0] = f_add(x[j0], comb8_mul(x[j1], tab));
        if constexpr (!INV) x[j1] = f_add(x[j1], x[j0]);
      }
    }
  }

  if (active) {
#pragma unroll
    for (int j = 0; j < (1 << K); ++j) dst[base + (uint64_t)j * stride] = x[j];
  }
  }
}

// Plain replication (used only when no layer is left after the copy layers).
__global__ void replicate_kernel(const uint4* __restrict__ msg, uint4* __restrict__ cw, uint64_t n_msg, uint64_t n_cw) {
  const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_cw) cw[i] = msg[i & (n_msg - 1)];
}

constexpr int NTT_KMAX = 4;
// Default layers per pass: 2^3 keeps ~100 VGPRs (4 waves/SIMD) and beat 2^4 on MI355X.
constexpr int NTT_KDEFAULT = 3;

// Layers per pass (the radix-2^K of a pass).  FLOCK_NTT_RADIX_LOG (1..4)
// overrides the default for schedule sweeps; read once per process.
static int ntt_radix_cap() {
  static const int cap = [] {
    const char* v = getenv("FLOCK_NTT_RADIX_LOG");
    int k = v ? atoi(v) : NTT_KDEFAULT;
    return (k >= 1 && k <= NTT_KMAX) ? k : NTT_KMAX;
  }();
  return cap;
}

// FLOCK_NTT_SHARED_TABLES=0 keeps every pass on the per-group nibble-comb kernel.
static bool ntt_shared_tables_enabled() {
  static const bool on = [] {
    const char* v = getenv("FLOCK_NTT_SHARED_TABLES");
    return !(v && v[0] == '0');
  }();
  return on;
}

// Product kernels: FLOCK_NTT_COMB=0|1|2 selects the comb of the shared-table
// passes (see ntt_pass_shared_kernel), FLOCK_NTT_PAIR_TABLES=0 the 4-bit Horner
// comb for the per-group passes (both for A/B runs).
static bool env_off(const char* name) {
  const char* v = getenv(name);
  return v && v[0] == '0';
}
static int ntt_comb() {
  static const int c = [] {
    const char* v = getenv("FLOCK_NTT_COMB");
    return (v && v[0] >= '0' && v[0] <= '2') ? v[0] - '0' : 2;
  }();
  return c;
}
// Sets per thread in the shared-table passes (FLOCK_NTT_SETS = 1, 2 or 4):
// a thread runs its sets one after the other on the work-group's tables, so a
// table build serves twice the products.  Measured on the 2^17 x 64 encode:
// 2 sets 0.396-0.404 ms, 1 set 0.406-0.408 ms, 4 sets no better than 2.
static int ntt_sets_per_thread() {
  static const int n = [] {
    const char* v = getenv("FLOCK_NTT_SETS");
    const int k = v ? atoi(v) : 2;
    return (k == 2 || k == 4) ? k : 1;
  }();
  return n;
}
static bool ntt_pair_tables_enabled() {
  static const bool on = !env_off("FLOCK_NTT_PAIR_TABLES");
  return on;
}
// FLOCK_NTT_POS_TABLES=0 keeps the deepest passes on the nibble-pair kernel.
static bool ntt_pos_tables_enabled() {
  static const bool on = !env_off("FLOCK_NTT_POS_TABLES");
  return on;
}

template <int K>
static int launch_pass(const uint4* src, uint4* dst, const uint4* tw, uint32_t log_d, uint32_t L, uint32_t log_lanes,
                       uint64_t src_pos_mask, uint32_t log_group, bool inverse, hipStream_t st) {
  const uint64_t n_threads = (1ull << (log_d - K)) << log_lanes;
  const uint64_t blocks = (n_threads + NTT_WG - 1) / NTT_WG;
  // Sets per work-group = NTT_WG / num_ntts; they share the block prefix when the
  // low position bits of the pass cover them.
  const uint32_t lo_bits = log_d - L - K;
  const int log_sets_per_wg = 8 - (int)log_lanes;
  if (ntt_shared_tables_enabled() && n_threads >= (uint64_t)NTT_WG &&
      (log_sets_per_wg <= 0 || (int)lo_bits >= log_sets_per_wg)) {
    const int comb = ntt_comb();
    const int sets = ntt_sets_per_thread();
    const int log_sets = sets == 4 ? 2 : (sets == 2 ? 1 : 0);
    if (!inverse && comb == 2 && log_sets > 0 && n_threads >= ((uint64_t)NTT_WG << log_sets) &&
        (log_sets_per_wg + log_sets <= 0 || (int)lo_bits >= log_sets_per_wg + log_sets)) {
      const uint64_t blocks_s = (n_threads + ((uint64_t)NTT_WG << log_sets) - 1) >> (8 + log_sets);
      if (log_sets == 2)
        hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 2, false, 4>), dim3((unsigned)blocks_s), dim3(NTT_WG), 0, st,
                           src, dst, tw, log_d, L, log_lanes, src_pos_mask, n_threads);
      else
        hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 2, false, 2>), dim3((unsigned)blocks_s), dim3(NTT_WG), 0, st,
                           src, dst, tw, log_d, L, log_lanes, src_pos_mask, n_threads);
      FLK_LAUNCH_CHECK("ntt_pass_shared_kernel");
      return FLK_OK;
    }
    if (inverse)  // the inverse runs the default product kernels only
      hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 2, true>), dim3((unsigned)blocks), dim3(NTT_WG), 0, st, src, dst,
                         tw, log_d, L, log_lanes, src_pos_mask, n_threads);
    else if (comb == 2)
      hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 2>), dim3((unsigned)blocks), dim3(NTT_WG), 0, st, src, dst, tw,
                         log_d, L, log_lanes, src_pos_mask, n_threads);
    else if (comb == 1)
      hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 1>), dim3((unsigned)blocks), dim3(NTT_WG), 0, st, src, dst, tw,
                         log_d, L, log_lanes, src_pos_mask, n_threads);
    else
      hipLaunchKernelGGL((ntt_pass_shared_kernel<K, 0>), dim3((unsigned)blocks), dim3(NTT_WG), 0, st, src, dst, tw,
                         log_d, L, log_lanes, src_pos_mask, n_threads);
    FLK_LAUNCH_CHECK("ntt_pass_shared_kernel");
    return FLK_OK;
  }
  if constexpr (K <= 3) {
    // One set per wave (num_ntts >= 64): per-wave nibble-position tables.
    if (log_group == 6 && ntt_pos_tables_enabled()) {
      const unsigned pos_blocks = (unsigned)((n_threads + NTT_POS_WG - 1) / NTT_POS_WG);
      if (inverse)
        hipLaunchKernelGGL((ntt_pass_pos_kernel<K, true>), dim3(pos_blocks), dim3(NTT_POS_WG), 0, st, src, dst, tw,
                           log_d, L, log_lanes, src_pos_mask, n_threads);
      else
        hipLaunchKernelGGL((ntt_pass_pos_kernel<K>), dim3(pos_blocks), dim3(NTT_POS_WG), 0, st, src, dst, tw, log_d, L,
                           log_lanes, src_pos_mask, n_threads);
      FLK_LAUNCH_CHECK("ntt_pass_pos_kernel");
      return FLK_OK;
    }
  }
  const bool pair = inverse || ntt_pair_tables_enabled();
  const size_t lds = (size_t)(NTT_WG / 64) * (64u >> log_group) * ((size_t)(pair ? 32 : 16) << (K - 1)) * sizeof(uint4);
  if (inverse)
    hipLaunchKernelGGL((ntt_pass_kernel<K, true, true>), dim3((unsigned)blocks), dim3(NTT_WG), lds, st, src, dst, tw,
                       log_d, L, log_lanes, src_pos_mask, n_threads, log_group);
  else if (pair)
    hipLaunchKernelGGL((ntt_pass_kernel<K, true>), dim3((unsigned)blocks), dim3(NTT_WG), lds, st, src, dst, tw, log_d, L,
                       log_lanes, src_pos_mask, n_threads, log_group);
  else
    hipLaunchKernelGGL((ntt_pass_kernel<K, false>), dim3((unsigned)blocks), dim3(NTT_WG), lds, st, src, dst, tw, log_d,
                       L, log_lanes, src_pos_mask, n_threads, log_group);
  FLK_LAUNCH_CHECK("ntt_pass_kernel");
  return FLK_OK;
}

// Pass schedule for layers [start, log_d): a list of layer counts (each 1..4).
// The deepest pass (whose blocks are too small to share twiddles across a
// work-group) takes up to 3 layers; the layers above it go in passes of 3 with
// the remainder as leading 2-layer passes (a remainder of 1 becomes 2 + 2).
// Measured on MI355X (2^17 x 64 encode): 2,2,3,3,3,3 0.409 ms, 1,3,3,3,3,3
// 0.433 ms, 4,3,3,3,3 0.412 ms (its 4-layer pass spills at 4 waves/SIMD).
// FLOCK_NTT_SCHEDULE="4,3,3,3,3" overrides it (must sum to the layer count).
// Schedules installed through flk_ntt_set_schedule (the autotuner's choices),
// keyed by (layers, log2 lanes); consulted after FLOCK_NTT_SCHEDULE and before
// the default rule.
static std::mutex g_sched_mu;
static std::map<std::pair<uint32_t, uint32_t>, std::vector<int>> g_sched;

static int ntt_schedule(uint32_t n, uint32_t log_lanes, int kmax, int* ks) {
  int cnt = 0;
  const char* env = getenv("FLOCK_NTT_SCHEDULE");
  if (env && env[0]) {
    uint32_t sum = 0;
    for (const char* p = env; *p;) {
      const int k = atoi(p);
      if (k < 1 || k > NTT_KMAX || cnt >= 64) return -1;
      ks[cnt++] = k;
      sum += (uint32_t)k;
      while (*p && *p != ',') ++p;
      if (*p == ',') ++p;
    }
    if (sum == n) return cnt;
    cnt = 0;  // does not match this transform: fall back to the default
  }
  {
    std::lock_guard<std::mutex> lk(g_sched_mu);
    auto it = g_sched.find({n, log_lanes});
    if (it != g_sched.end()) {
      for (int k : it->second) ks[cnt++] = k;
      return cnt;
    }
  }
  if (n == 0) return 0;
  const uint32_t deep = n < (uint32_t)kmax ? n : (uint32_t)kmax;
  uint32_t rest = n - deep;
  if (kmax >= 3 && rest % 3 == 1 && rest >= 4) {
    // 2 + 2 rather than 1 + 3: the first passes are bound by HBM (the first also
    // reads the replicated message), so a layer moved into them is nearly free
    // while the 3-layer passes are bound by table lookups.
    ks[cnt++] = 2;
    ks[cnt++] = 2;
    rest -= 4;
  } else if (kmax >= 3 && rest % 3 != 0) {
    ks[cnt++] = (int)(rest % 3);
    rest -= rest % 3;
  }
  while (rest > 0) {
    const uint32_t k = rest < (uint32_t)kmax ? rest : (uint32_t)kmax;
    ks[cnt++] = (int)k;
    rest -= k;
  }
  ks[cnt++] = (int)deep;
  return cnt;
}

// Run layers [start, log_d) reading the first pass from `src` (positions masked
// by src_pos_mask) and writing all passes to `dst`.  inverse: the same passes in
// reverse order, each running its layers deepest first (the inverse transform,
// CF/ntt/additive_ntt_f128.rs:6322-6342).
static int run_layers(const uint4* src, uint4* dst, const uint4* tw, uint32_t log_d, uint32_t start,
                      uint32_t log_lanes, uint64_t src_pos_mask, hipStream_t st, bool inverse = false) {
  const uint32_t log_group = log_lanes < 6 ? log_lanes : 6;
  // The per-group kernel holds (64/G) * 2^(K-1) tables of 512 B per wave: K <= log2(G) + 1.
  const int kgroup = (int)log_group + 1;
  int kmax = ntt_radix_cap();
  int ks[64];
  const int n_pass = ntt_schedule(log_d - start, log_lanes, kmax, ks);
  if (n_pass < 0) return set_error(FLK_E_SHAPE, "FLOCK_NTT_SCHEDULE is malformed");
  // Launch plan in forward order: (first layer, layers), splitting passes the
  // per-group kernel cannot hold.
  uint32_t plan_L[64 * 4];
  int plan_k[64 * 4];
  int n_plan = 0;
  uint32_t L = start;
  for (int pi = 0; pi < n_pass; ++pi) {
    const int K = ks[pi];
    const uint32_t lo_bits = log_d - L - (uint32_t)K;
    const int log_sets_per_wg = 8 - (int)log_lanes;
    const bool shared = ntt_shared_tables_enabled() && ((log_d - (uint32_t)K + log_lanes) >= 8) &&
                        (log_sets_per_wg <= 0 || (int)lo_bits >= log_sets_per_wg);
    int todo = K;
    while (todo > 0) {
      const int k = shared ? todo : (todo < kgroup ? todo : kgroup);
      plan_L[n_plan] = L;
      plan_k[n_plan++] = k;
      L += (uint32_t)k;
      todo -= k;
    }
  }
  const uint4* s = src;
  uint64_t mask = src_pos_mask;
  for (int q = 0; q < n_plan; ++q) {
    const int p = inverse ? n_plan - 1 - q : q;
    int rc;
    switch (plan_k[p]) {
      case 1: rc = launch_pass<1>(s, dst, tw, log_d, plan_L[p], log_lanes, mask, log_group, inverse, st); break;
      case 2: rc = launch_pass<2>(s, dst, tw, log_d, plan_L[p], log_lanes, mask, log_group, inverse, st); break;
      case 3: rc = launch_pass<3>(s, dst, tw, log_d, plan_L[p], log_lanes, mask, log_group, inverse, st); break;
      default: rc = launch_pass<4>(s, dst, tw, log_d, plan_L[p], log_lanes, mask, log_group, inverse, st); break;
    }
    if (rc) return rc;
    s = dst;
    mask = ~0ull;
  }
  return FLK_OK;
}

int ntt_rs_encode_device(size_t dim, const uint4* msg, size_t n_msg, uint4* cw, size_t n_cw, size_t num_ntts,
                         hipStream_t st) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!msg || !cw) return set_error(FLK_E_SHAPE, "rs_encode: null pointer");
  if (!is_pow2(num_ntts) || n_msg == 0 || n_msg % num_ntts || n_cw % n_msg)
    return set_error(FLK_E_SHAPE, "rs_encode: bad shape n_msg=%zu n_cw=%zu num_ntts=%zu", n_msg, n_cw, num_ntts);
  const size_t inv_rate = n_cw / n_msg;
  const int log_inv_rate = log2_exact(inv_rate);
  const int log_d = log2_exact(n_cw / num_ntts);
  const int log_lanes = log2_exact(num_ntts);
  if (log_inv_rate < 1 || log_d < 0 || log_inv_rate > log_d || (size_t)log_d > dim)
    return set_error(FLK_E_SHAPE, "rs_encode: bad rate/domain (inv_rate=%zu log_d=%d dim=%zu)", inv_rate, log_d, dim);
  if (dim > 24) return set_error(FLK_E_SHAPE, "rs_encode: dim %zu exceeds the device twiddle cache (24)", dim);
  if ((const void*)msg < (const void*)(cw + n_cw) && (const void*)cw < (const void*)(msg + n_msg))
    return set_error(FLK_E_SHAPE, "rs_encode: msg and codeword overlap");
  const uint4* tw = twiddles_device(dim, &rc);
  if (rc) return rc;
  const uint64_t msg_pos_mask = (1ull << (log_d - log_inv_rate)) - 1ull;
  if (log_inv_rate == log_d) {
    const uint64_t blocks = (n_cw + 255) / 256;
    hipLaunchKernelGGL(replicate_kernel, dim3((unsigned)blocks), dim3(256), 0, st, msg, cw, (uint64_t)n_msg,
                       (uint64_t)n_cw);
    FLK_LAUNCH_CHECK("replicate_kernel");
    return FLK_OK;
  }
  return run_layers(msg, cw, tw, (uint32_t)log_d, (uint32_t)log_inv_rate, (uint32_t)log_lanes, msg_pos_mask, st);
}

}  // namespace flk

extern "C" {

int flk_ntt_rs_encode_interleaved(size_t dim, const flk_f128* msg, size_t n_msg, flk_f128* codeword, size_t n_cw,
                                  size_t num_ntts, void* stream) {
  return flk::ntt_rs_encode_device(dim, reinterpret_cast<const uint4*>(msg), n_msg, reinterpret_cast<uint4*>(codeword),
                                   n_cw, num_ntts, static_cast<hipStream_t>(stream));
}

int flk_ntt_forward_interleaved(size_t dim, flk_f128* data, size_t n_total, size_t num_ntts, size_t start_layer,
                                void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!data) return flk::set_error(FLK_E_SHAPE, "forward_interleaved: null pointer");
  if (!flk::is_pow2(num_ntts) || n_total % num_ntts)
    return flk::set_error(FLK_E_SHAPE, "forward_interleaved: bad shape n_total=%zu num_ntts=%zu", n_total, num_ntts);
  const int log_d = flk::log2_exact(n_total / num_ntts);
  if (log_d < 0 || (size_t)log_d > dim || start_layer > (size_t)log_d)
    return flk::set_error(FLK_E_SHAPE, "forward_interleaved: bad domain log_d=%d dim=%zu start=%zu", log_d, dim,
                          start_layer);
  if (dim > 24) return flk::set_error(FLK_E_SHAPE, "forward_interleaved: dim %zu exceeds 24", dim);
  if ((size_t)log_d == start_layer) return FLK_OK;
  const uint4* tw = flk::twiddles_device(dim, &rc);
  if (rc) return rc;
  uint4* d = reinterpret_cast<uint4*>(data);
  return flk::run_layers(d, d, tw, (uint32_t)log_d, (uint32_t)start_layer, (uint32_t)flk::log2_exact(num_ntts), ~0ull,
                         static_cast<hipStream_t>(stream));
}

int flk_ntt_inverse_interleaved(size_t dim, flk_f128* data, size_t n_total, size_t num_ntts, void* stream) {
  int rc = flk::check_enabled();
  if (rc) return rc;
  if (!data) return flk::set_error(FLK_E_SHAPE, "inverse_interleaved: null pointer");
  if (!flk::is_pow2(num_ntts) || n_total % num_ntts)
    return flk::set_error(FLK_E_SHAPE, "inverse_interleaved: bad shape n_total=%zu num_ntts=%zu", n_total, num_ntts);
  const int log_d = flk::log2_exact(n_total / num_ntts);
  if (log_d < 0 || (size_t)log_d > dim)
    return flk::set_error(FLK_E_SHAPE, "inverse_interleaved: bad domain log_d=%d dim=%zu", log_d, dim);
  if (dim > 24) return flk::set_error(FLK_E_SHAPE, "inverse_interleaved: dim %zu exceeds 24", dim);
  if (log_d == 0) return FLK_OK;
  const uint4* tw = flk::twiddles_device(dim, &rc);
  if (rc) return rc;
  uint4* d = reinterpret_cast<uint4*>(data);
  return flk::run_layers(d, d, tw, (uint32_t)log_d, 0, (uint32_t)flk::log2_exact(num_ntts), ~0ull,
                         static_cast<hipStream_t>(stream), true);
}

int flk_ntt_set_schedule(size_t n_layers, size_t log_lanes, const int* ks, size_t n) {
  if (n_layers == 0 || n_layers > 40 || log_lanes > 16)
    return flk::set_error(FLK_E_SHAPE, "ntt_set_schedule: n_layers=%zu log_lanes=%zu", n_layers, log_lanes);
  std::lock_guard<std::mutex> lk(flk::g_sched_mu);
  const std::pair<uint32_t, uint32_t> key((uint32_t)n_layers, (uint32_t)log_lanes);
  if (n == 0) {
    flk::g_sched.erase(key);
    return FLK_OK;
  }
  if (!ks || n > 40) return flk::set_error(FLK_E_SHAPE, "ntt_set_schedule: null or too long schedule");
  size_t sum = 0;
  for (size_t i = 0; i < n; ++i) {
    if (ks[i] < 1 || ks[i] > flk::NTT_KMAX)
      return flk::set_error(FLK_E_SHAPE, "ntt_set_schedule: pass %zu has %d layers (1..%d)", i, ks[i], flk::NTT_KMAX);
    sum += (size_t)ks[i];
  }
  if (sum != n_layers) return flk::set_error(FLK_E_SHAPE, "ntt_set_schedule: passes sum to %zu, not %zu", sum, n_layers);
  flk::g_sched[key] = std::vector<int>(ks, ks + n);
  return FLK_OK;
}

int flk_ntt_get_schedule(size_t n_layers, size_t log_lanes, int* ks, size_t cap, size_t* n) {
  if (!ks || !n) return flk::set_error(FLK_E_SHAPE, "ntt_get_schedule: null pointer");
  if (n_layers > 40 || log_lanes > 16)
    return flk::set_error(FLK_E_SHAPE, "ntt_get_schedule: n_layers=%zu log_lanes=%zu", n_layers, log_lanes);
  int tmp[64];
  const int cnt = flk::ntt_schedule((uint32_t)n_layers, (uint32_t)log_lanes, flk::ntt_radix_cap(), tmp);
  if (cnt < 0) return flk::set_error(FLK_E_SHAPE, "FLOCK_NTT_SCHEDULE is malformed");
  if ((size_t)cnt > cap) return flk::set_error(FLK_E_SHAPE, "ntt_get_schedule: %d passes, capacity %zu", cnt, cap);
  for (int i = 0; i < cnt; ++i) ks[i] = tmp[i];
  *n = (size_t)cnt;
  return FLK_OK;
}

}  // extern "C"
