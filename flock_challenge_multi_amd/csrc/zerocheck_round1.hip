// Zerocheck round 1 (univariate skip over k_skip = 6) for gfx950.
//
// Semantics: round1_naive (CF/zerocheck/univariate_skip.rs:90-152) with the
// protocol's challenge layout (CF/zerocheck.rs:1680-1692): r[6..9) are the
// small challenges phi8(0xF7, 0x53, 0xB5), r[9..13) the medium challenges
// x^(2^i) / (1 + x^(2^i)), r[13..m) = r_outer
// (CF/zerocheck/univariate_skip_optimized.rs:136-232).  For that layout the
// eq weight of row x = (K, J, x_outer) factors as
//     eq(x) = C_s alpha^K * x^J / D * eq(r_outer, x_outer)
// with alpha = phi8(0x02), C_s = phi8(0x1C), D = prod_{i<4} (1 + x^(2^i)), so
//     P[l] = (C_s / D) sum_{x_outer} eq_outer[x_outer] *
//            sum_J x^J phi8( sum_K alpha^K v_l(K, J, x_outer) )
// and the inner sums are GF(2^8) work (Horner in alpha = xtime on packed bytes)
// plus bit shifts by J; one GF(2^128) product per (x_outer, l).
//
// Work-group = one 128-row block (one x_outer) at a time:
//   phase 1, lane = row (K = lane & 7, J = lane >> 3): the 64-point LDE of the
//     row's a, b, c bits (S -> Lambda: a fixed GF(2)-linear map 64 bits -> 64
//     bytes, applied with 16 nibble tables of 64 bytes), v_ab = a_l * b_l in
//     GF(2^8) by log/exp tables, v_c = c_l; both stored to LDS;
//   phase 2, lane = (byte word w, which, J-group): Horner over K for 4 J values
//     and 4 l at once (packed bytes), phi8 lookups shifted by J, a 4-lane XOR
//     shuffle to finish the J sum, then one byte-comb product by eq_outer.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

struct Round1Tables {
  uint32_t lde[16 * 16 * 16];  // [nibble pos p][value v][word w]: bytes l = 4w..4w+3
  uint16_t logt[256];          // log_3 of v; log(0) = 511 (lands in the zero tail of exp)
  uint8_t expt[1024];          // 3^i for i < 510, zero beyond
  uint4 phi[256];              // phi8 table
};

// ---- host GF(2^8) (AES polynomial) and its additive NTT (CF/ntt.rs:22-169) ----
static uint8_t g8mul(uint8_t a, uint8_t b) {
  uint16_t acc = 0;
  for (int i = 0; i < 8; ++i)
    if ((a >> i) & 1) acc ^= (uint16_t)((uint16_t)b << i);
  for (int i = 15; i >= 8; --i)
    if ((acc >> i) & 1) acc ^= (uint16_t)(0x11B << (i - 8));
  return (uint8_t)acc;
}
static uint8_t g8inv(uint8_t a) {
  uint8_t r = 1, s = a;
  for (int i = 0; i < 8; ++i) {
    if ((0xFE >> i) & 1) r = g8mul(r, s);
    s = g8mul(s, s);
  }
  return r;
}
static void g8_twiddles(int k, uint8_t beta, uint8_t* tw) {
  size_t len = size_t(1) << (k - 1);
  uint8_t layer[128];
  for (size_t i = 0; i < len; ++i) layer[i] = beta ^ (uint8_t)(2 * i);
  uint8_t root = 1;
  size_t at = len;
  for (size_t i = 0; i < len; ++i) tw[at - 1 + i] = layer[i];
  for (int it = 1; it < k; ++it) {
    at >>= 1;
    const uint8_t s01 = layer[1] ^ layer[0];
    const uint8_t next_root = g8mul(s01, s01) ^ g8mul(root, s01);
    for (size_t i = 0; i < at; ++i) layer[i] = g8mul(layer[2 * i], layer[2 * i]) ^ g8mul(root, layer[2 * i]);
    len = at;
    root = next_root;
    const uint8_t sinv = g8inv(root);
    for (size_t j = 0; j < len; ++j) tw[at - 1 + j] = g8mul(sinv, layer[j]);
  }
}
static void g8_fft(uint8_t* v, size_t n, const uint8_t* tw, size_t idx) {
  if (n == 1) return;
  const size_t h = n / 2;
  for (size_t i = 0; i < h; ++i) {
    const uint8_t w = v[h + i];
    v[i] ^= g8mul(tw[idx - 1], w);
    v[h + i] = w ^ v[i];
  }
  g8_fft(v, h, tw, 2 * idx);
  g8_fft(v + h, h, tw, 2 * idx + 1);
}
static void g8_ifft(uint8_t* v, size_t n, const uint8_t* tw, size_t idx) {
  if (n == 1) return;
  const size_t h = n / 2;
  g8_ifft(v, h, tw, 2 * idx);
  g8_ifft(v + h, h, tw, 2 * idx + 1);
  for (size_t i = 0; i < h; ++i) {
    v[h + i] ^= v[i];
    v[i] ^= g8mul(tw[idx - 1], v[h + i]);
  }
}

static flk_f128 phi8_host(uint8_t v) {
  const flk_f128 g{0x6b8330483c2e9849ull, 0x0dcb364640a222feull};  // phi8(0x02), CF/field/phi8.rs:28-31
  flk_f128 acc{0, 0}, p{1, 0};
  for (int i = 0; i < 8; ++i) {
    if ((v >> i) & 1) acc = flk_f128{acc.lo ^ p.lo, acc.hi ^ p.hi};
    p = host_mul(p, g);
  }
  return acc;
}

static void build_round1_tables(Round1Tables& T) {
  uint8_t twS[64], twL[64];
  g8_twiddles(6, 0, twS);
  g8_twiddles(6, 64, twL);
  uint8_t M[64][64];  // M[s][l] = LDE of the unit vector e_s at Lambda point l
  for (int s = 0; s < 64; ++s) {
    uint8_t v[64] = {0};
    v[s] = 1;
    g8_ifft(v, 64, twS, 1);
    g8_fft(v, 64, twL, 1);
    memcpy(M[s], v, 64);
  }
  for (int p = 0; p < 16; ++p)
    for (int val = 0; val < 16; ++val) {
      uint8_t bytes[64] = {0};
      for (int j = 0; j < 4; ++j)
        if ((val >> j) & 1)
          for (int l = 0; l < 64; ++l) bytes[l] ^= M[4 * p + j][l];
      memcpy(&T.lde[(p * 16 + val) * 16], bytes, 64);
    }
  uint8_t e = 1;
  for (int i = 0; i < 255; ++i) {
    T.expt[i] = e;
    T.expt[i + 255] = e;
    T.logt[e] = (uint16_t)i;
    e = g8mul(e, 3);
  }
  for (int i = 510; i < 1024; ++i) T.expt[i] = 0;
  T.logt[0] = 511;
  for (int v = 0; v < 256; ++v) {
    const flk_f128 f = phi8_host((uint8_t)v);
    T.phi[v] = make_uint4((uint32_t)f.lo, (uint32_t)(f.lo >> 32), (uint32_t)f.hi, (uint32_t)(f.hi >> 32));
  }
}

// C_s / D with C_s = phi8(0x1C), D = prod_{i<4} (1 + x^(2^i)).
static flk_f128 round1_scale() {
  flk_f128 d{1, 0};
  for (int i = 0; i < 4; ++i) d = host_mul(d, flk_f128{1ull ^ (1ull << (1u << i)), 0});
  return host_mul(phi8_host(0x1C), host_inv(d));
}

static std::mutex g_r1_mu;
static std::vector<std::pair<int, Round1Tables*>> g_r1_dev;

static const Round1Tables* round1_tables_device(int* status) {
  *status = FLK_OK;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    *status = device_failure(e, "hipGetDevice");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_r1_mu);
  for (auto& kv : g_r1_dev)
    if (kv.first == dev) return kv.second;
  static Round1Tables host;
  static bool built = false;
  if (!built) {
    build_round1_tables(host);
    built = true;
  }
  Round1Tables* d = nullptr;
  e = hipMalloc(&d, sizeof(Round1Tables));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *status = set_error(FLK_E_NOMEM, "round1 tables: hipMalloc failed");
    return nullptr;
  }
  e = hipMemcpy(d, &host, sizeof(Round1Tables), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d);
    *status = device_failure(e, "round1 tables upload");
    return nullptr;
  }
  g_r1_dev.emplace_back(dev, d);
  return d;
}

constexpr int R1_WG = 128;

__device__ __forceinline__ void atomic_xor4_r1(uint4* out, uint4 v) {
  uint32_t* o = reinterpret_cast<uint32_t*>(out);
  atomicXor(o + 0, v.x);
  atomicXor(o + 1, v.y);
  atomicXor(o + 2, v.z);
  atomicXor(o + 3, v.w);
}

__device__ __forceinline__ uint32_t xtime4(uint32_t w) {
  return ((w & 0x7F7F7F7Fu) << 1) ^ (((w >> 7) & 0x01010101u) * 0x1Bu);
}

__global__ void __launch_bounds__(R1_WG) zc_round1_kernel(const uint64_t* __restrict__ A, const uint64_t* __restrict__ B,
                                                         const uint64_t* __restrict__ C, uint64_t n_blocks,
                                                         const uint4* __restrict__ eq_outer,
                                                         const Round1Tables* __restrict__ T, uint4* __restrict__ out) {
  __shared__ uint4 lde4[16 * 16 * 4];  // 16 KiB
  __shared__ uint16_t logt[256];
  __shared__ uint8_t expt[1024];
  __shared__ uint4 phi[256];
  __shared__ uint32_t rowdata[R1_WG * 33];  // [row][ab words 0..15 | c words 16..31], padded
  __shared__ uint4 nib[32];
  __shared__ uint4 etab[256];
  __shared__ uint32_t ypart[2 * 16 * 4 * 4 * 5];
  const uint32_t t = threadIdx.x;
  {
    const uint4* src = reinterpret_cast<const uint4*>(T->lde);
    for (uint32_t i = t; i < 16 * 16 * 4; i += R1_WG) lde4[i] = src[i];
    for (uint32_t i = t; i < 256; i += R1_WG) {
      logt[i] = T->logt[i];
      phi[i] = T->phi[i];
    }
    for (uint32_t i = t; i < 1024; i += R1_WG) expt[i] = T->expt[i];
  }
  // phase-2a role: 16 words x 2 (ab, c) x 4 J-groups
  const uint32_t w = t & 15u, which = (t >> 4) & 1u, jg = t >> 5;
  U256 acc;
  u256_zero(acc);
  __syncthreads();

  for (uint64_t blk = blockIdx.x; blk < n_blocks; blk += gridDim.x) {
    // ---- phase 1: one row per lane ----
    const uint64_t row = blk * R1_WG + t;
    const uint64_t a = A[row], b = B[row], c = C[row];
    uint32_t ah[16], bh[16], ch[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) ah[q] = bh[q] = ch[q] = 0u;
#pragma unroll
    for (int p = 0; p < 16; ++p) {
      const uint32_t na = (uint32_t)(a >> (4 * p)) & 15u, nb = (uint32_t)(b >> (4 * p)) & 15u,
                     nc = (uint32_t)(c >> (4 * p)) & 15u;
      const uint4* ea = lde4 + (p * 16 + na) * 4;
      const uint4* eb = lde4 + (p * 16 + nb) * 4;
      const uint4* ec = lde4 + (p * 16 + nc) * 4;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4 x = ea[q], y = eb[q], z = ec[q];
        ah[4 * q] ^= x.x; ah[4 * q + 1] ^= x.y; ah[4 * q + 2] ^= x.z; ah[4 * q + 3] ^= x.w;
        bh[4 * q] ^= y.x; bh[4 * q + 1] ^= y.y; bh[4 * q + 2] ^= y.z; bh[4 * q + 3] ^= y.w;
        ch[4 * q] ^= z.x; ch[4 * q + 1] ^= z.y; ch[4 * q + 2] ^= z.z; ch[4 * q + 3] ^= z.w;
      }
    }
    uint32_t* rd = rowdata + t * 33;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      uint32_t v = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t x = (ah[q] >> (8 * j)) & 255u, y = (bh[q] >> (8 * j)) & 255u;
        v |= (uint32_t)expt[logt[x] + logt[y]] << (8 * j);
      }
      rd[q] = v;
      rd[16 + q] = ch[q];
    }
    // eq_outer[blk] byte-comb table (nibble tables of e and e*x^4, then bytes)
    if (t < 32) {
      uint4 e = eq_outer[blk];
      if (t >= 16) e = f_mul_x4(e);
      nib[t] = comb_entry(e, t & 15u);
    }
    __syncthreads();
    for (uint32_t i = t; i < 256; i += R1_WG) etab[i] = f_add(nib[i & 15u], nib[16 + (i >> 4)]);
    // ---- phase 2a: lane (w, which, jg): partial Y for 4 l and J in its group ----
    // Horner over K (alpha = 0x02 -> xtime) on packed bytes: 4 l per word.
    uint32_t wj[4];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const uint32_t J = 4 * jg + jj;
      uint32_t h = 0;
#pragma unroll
      for (int K = 7; K >= 0; --K) h = xtime4(h) ^ rowdata[(K + 8 * J) * 33 + which * 16 + w];
      wj[jj] = h;
    }
    uint32_t Y[4][5];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 5; ++q) Y[j][q] = 0u;
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      const uint32_t J = 4 * jg + jj;  // shift phi8(w) by J < 16 bits
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint4 f = phi[(wj[jj] >> (8 * j)) & 255u];
        if (J == 0) {
          Y[j][0] ^= f.x; Y[j][1] ^= f.y; Y[j][2] ^= f.z; Y[j][3] ^= f.w;
        } else {
          Y[j][0] ^= f.x << J;
          Y[j][1] ^= funnel(f.y, f.x, 32 - J);
          Y[j][2] ^= funnel(f.z, f.y, 32 - J);
          Y[j][3] ^= funnel(f.w, f.z, 32 - J);
          Y[j][4] ^= f.w >> (32 - J);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 5; ++q) ypart[((((which * 16 + w) * 4 + jg) * 4) + j) * 5 + q] = Y[j][q];
    __syncthreads();
    // ---- phase 2b: lane (l, which2): finish the J sum, multiply by eq_outer ----
    {
      const uint32_t l = t & 63u, wh = t >> 6, ww = l >> 2, j = l & 3u;
      U256 y;
      u256_zero(y);
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int q = 0; q < 5; ++q) y.w[q] ^= ypart[((((wh * 16 + ww) * 4 + g) * 4) + j) * 5 + q];
      comb8_mul_acc(f_reduce(y), etab, acc);
    }
    __syncthreads();
  }
  atomic_xor4_r1(out + t, f_reduce(acc));  // out[which * 64 + l], t = which * 64 + l
}

__global__ void scale_kernel(uint4* __restrict__ v, uint32_t n, uint4 c) {
  __shared__ uint4 tab[16];
  if (threadIdx.x < 16) tab[threadIdx.x] = comb_entry(c, threadIdx.x);
  __syncthreads();
  if (threadIdx.x < n) v[threadIdx.x] = comb_mul<1>(v[threadIdx.x], tab);
}

__global__ void zero128_kernel(uint4* p) { p[threadIdx.x] = make_uint4(0u, 0u, 0u, 0u); }

}  // namespace flk


extern "C" int flk_zc_round1(const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                             const flk_f128* r_outer, flk_f128* out_ab, flk_f128* out_c, void* stream) {
  using namespace flk;
  int rc = check_enabled();
  if (rc) return rc;
  if (!a_bits || !b_bits || !c_bits || !out_ab || !out_c || (m > 13 && !r_outer))
    return set_error(FLK_E_SHAPE, "zc_round1: null pointer");
  if (m < 13 || m > 40) return set_error(FLK_E_SHAPE, "zc_round1: m=%zu out of range (13..40)", m);
  if (((uintptr_t)a_bits | (uintptr_t)b_bits | (uintptr_t)c_bits) % 8)
    return set_error(FLK_E_SHAPE, "zc_round1: bit vectors must be 8-byte aligned");
  if (out_c != out_ab + 64) return set_error(FLK_E_SHAPE, "zc_round1: out_c must follow out_ab (one 128-entry buffer)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Round1Tables* T = round1_tables_device(&rc);
  if (rc) return rc;
  const size_t n_outer = size_t(1) << (m - 13);
  uint4* eq = nullptr;
  hipError_t e = hipMallocAsync(reinterpret_cast<void**>(&eq), n_outer * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "zc_round1: eq_outer allocation failed");
  }
  rc = flk_build_eq(r_outer, m - 13, reinterpret_cast<flk_f128*>(eq), stream);
  if (rc) return rc;
  uint4* out = reinterpret_cast<uint4*>(out_ab);
  hipLaunchKernelGGL(zero128_kernel, dim3(1), dim3(128), 0, st, out);
  FLK_LAUNCH_CHECK("zero128_kernel");
  // LDS (~53 KiB per work-group) allows 3 work-groups per CU: 768 fill the chip;
  // 2048 leave room for uneven block counts per work-group.
  uint64_t grid = n_outer < 2048 ? n_outer : 2048;
  hipLaunchKernelGGL(zc_round1_kernel, dim3((unsigned)grid), dim3(R1_WG), 0, st,
                     reinterpret_cast<const uint64_t*>(a_bits), reinterpret_cast<const uint64_t*>(b_bits),
                     reinterpret_cast<const uint64_t*>(c_bits), (uint64_t)n_outer, eq, T, out);
  FLK_LAUNCH_CHECK("zc_round1_kernel");
  const flk_f128 sc = round1_scale();
  hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(128), 0, st, out, 128u,
                     make_uint4((uint32_t)sc.lo, (uint32_t)(sc.lo >> 32), (uint32_t)sc.hi, (uint32_t)(sc.hi >> 32)));
  FLK_LAUNCH_CHECK("scale_kernel");
  e = hipFreeAsync(eq, st);
  if (e != hipSuccess) return device_failure(e, "zc_round1: free eq_outer");
  return FLK_OK;
}
