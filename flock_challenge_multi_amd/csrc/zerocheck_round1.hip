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
// and the inner sums are GF(2^8) work (Horner in alpha) plus x^J phi8 of the
// J-sums; one GF(2^128) product per (x_outer, l).
//
// P_C needs no per-row LDE: the LDE is GF(2^8)-linear and c is a bit, so
//     P_C[l] = sum_s phi8(M[s][l]) * Q[s],   Q[s] = sum_x eq(x) c_x[s],
// with M the 64 x 64 LDE matrix (S -> Lambda).  Q is a column sum of the bit
// matrix c: the ring-switch fold_1b_rows kernel computes it over row pairs (c
// read as 2^(m-7) elements of 128 bits, weighted by eq over r[7..m)), and a
// 64-thread epilogue binds r[6] and applies M.
//
// The P_AB kernel keeps the GF(2^8) work on bit planes: the LDE (a fixed
// GF(2)-linear map, 64 bits -> 64 bytes) is applied with nibble tables whose
// entries are already transposed, so a row's 64 values arrive as 8 planes per
// 32 points (plane p: bit p of each value).  A product of 32 value pairs is then
// 64 AND-XORs plus the reduction, with no lookups (a 64 KiB byte-product table
// costs a quarter of the CU's cycles in LDS bank conflicts), and the Horner step
// in alpha moves planes.  A lane owns the 8 rows (K, J) of one J of a block, so
// the sum over K never leaves it; the J-sums go through the wave's LDS as bytes.
// Every data-indexed 16-byte LDS table has 16 entries (conflict-free).
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

struct Round1Tables {
  uint32_t lde[16 * 16 * 16];  // [nibble pos p][value v][word w]: bit planes of the LDE image (below)
  uint8_t M[64 * 64];          // [s][l]: the LDE of the unit vector e_s at point l
  uint4 phi_pow[8];            // phi8(2^i)
  uint4 xj[16 * 32];           // [J][n] = x^J * phi8(n), [J][16 + n] = x^J * phi8(n << 4), reduced
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
      // bit planes: word w = 8h + p' holds bit p' of bytes l = 32h .. 32h + 31
      for (int w = 0; w < 16; ++w) {
        uint32_t word = 0;
        for (int i = 0; i < 32; ++i) word |= (uint32_t)((bytes[32 * (w >> 3) + i] >> (w & 7)) & 1) << i;
        T.lde[(p * 16 + val) * 16 + w] = word;
      }
    }
  memcpy(T.M, M, sizeof(T.M));
  auto as_u4 = [](flk_f128 f) {
    return make_uint4((uint32_t)f.lo, (uint32_t)(f.lo >> 32), (uint32_t)f.hi, (uint32_t)(f.hi >> 32));
  };
  for (int i = 0; i < 8; ++i) T.phi_pow[i] = as_u4(phi8_host((uint8_t)(1u << i)));
  for (int J = 0; J < 16; ++J) {
    const flk_f128 xJ{1ull << J, 0};
    for (int n = 0; n < 16; ++n) {
      T.xj[J * 32 + n] = as_u4(host_mul(xJ, phi8_host((uint8_t)n)));
      T.xj[J * 32 + 16 + n] = as_u4(host_mul(xJ, phi8_host((uint8_t)(n << 4))));
    }
  }
}

// C_s / D with C_s = phi8(0x1C), D = prod_{i<4} (1 + x^(2^i)).
static inline uint4 u4_of(flk_f128 a) {
  return make_uint4((uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32));
}

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

constexpr int R1_ROWS = 128;   // rows of one block (one x_outer)
constexpr int R1_NW = 8;       // waves per work-group
constexpr int R1_QUAD = 4;     // blocks per wave pass: lane = (block b, J)
constexpr uint32_t R1_XS = 17; // words per (b, J) row of a wave's exchange (16 + 1: conflict-free stores)

// GF(2^8) (x^8 + x^4 + x^3 + x + 1) products of 32 pairs at once on bit planes:
// c[p] bit i = bit p of a_i * b_i, where a[p] bit i = bit p of a_i.  Schoolbook
// (an AND-XOR per term in one V_BITOP3_B32, truth table 0x6A = (s0 & s1) ^ s2),
// then x^k = x^(k-4) + x^(k-5) + x^(k-7) + x^(k-8) for k = 14 .. 8.
__device__ __forceinline__ void gf8_mul_planes(const uint32_t* a, const uint32_t* b, uint32_t* c) {
  uint32_t t[15];
#pragma unroll
  for (int k = 0; k < 15; ++k) {
    const int i0 = k < 8 ? 0 : k - 7, i1 = k < 8 ? k : 7;
    t[k] = a[i0] & b[k - i0];
#pragma unroll
    for (int i = i0 + 1; i <= i1; ++i) t[k] = __builtin_amdgcn_bitop3_b32(a[i], b[k - i], t[k], 0x6A);
  }
  // t12..t14 fold into t8..t10 first (x^12 = x^8 + x^7 + x^5 + x^4, ...), then
  // t8..t14 into t0..t7, with three-input XORs
  const uint32_t t8 = xor3(t[8], t[12], t[13]), t9 = xor3(t[9], t[13], t[14]), t10 = t[10] ^ t[14];
  const uint32_t t11 = t[11], t12 = t[12], t13 = t[13], t14 = t[14];
  c[0] = t[0] ^ t8;
  c[1] = xor3(t[1], t8, t9);
  c[2] = xor3(t[2], t9, t10);
  c[3] = xor3(t[3], t8, t10) ^ t11;
  c[4] = xor3(xor3(t[4], t8, t9), t11, t12);
  c[5] = xor3(xor3(t[5], t9, t10), t12, t13);
  c[6] = xor3(xor3(t[6], t10, t11), t13, t14);
  c[7] = xor3(t[7], t11, t12) ^ t14;
}

// u <- x * u + v on bit planes (x = alpha = 0x02): the planes move up one, plane
// 7 folds into planes 0, 1, 3, 4.
__device__ __forceinline__ void gf8_horner_planes(uint32_t* u, const uint32_t* v) {
  const uint32_t u7 = u[7];
  u[7] = u[6] ^ v[7];
  u[6] = u[5] ^ v[6];
  u[5] = u[4] ^ v[5];
  u[4] = xor3(u[3], u7, v[4]);
  u[3] = xor3(u[2], u7, v[3]);
  u[2] = u[1] ^ v[2];
  u[1] = xor3(u[0], u7, v[1]);
  u[0] = u7 ^ v[0];
}

// Bit planes -> bytes: P[p] bit i = bit p of byte i (i < 32) becomes w[q] byte j =
// byte 4q + j.  Three delta-swap stages transpose the 8 x 8 bit blocks of each
// byte column (afterwards P[c] byte k = byte 8k + c), then byte permutes gather
// w[2k + g] = (P[4g] .. P[4g + 3]) byte k.
__device__ __forceinline__ void planes_to_bytes(uint32_t* P, uint32_t* w) {
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const int d = 1 << s;
    const uint32_t mask = s == 0 ? 0x55555555u : (s == 1 ? 0x33333333u : 0x0F0F0F0Fu);
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      if (p & d) continue;
      const uint32_t x = __builtin_amdgcn_bitop3_b32(P[p] >> d, P[p + d], mask, 0x28);  // (s0 ^ s1) & s2
      P[p + d] ^= x;
      P[p] ^= x << d;
    }
  }
#pragma unroll
  for (int g = 0; g < 2; ++g) {
    // x01 = (P0.b0, P1.b0, P0.b2, P1.b2), y01 = (P0.b1, P1.b1, P0.b3, P1.b3), likewise x23, y23
    const uint32_t* Q = P + 4 * g;
    const uint32_t x01 = __builtin_amdgcn_perm(Q[1], Q[0], 0x06020400u), y01 = __builtin_amdgcn_perm(Q[1], Q[0], 0x07030501u);
    const uint32_t x23 = __builtin_amdgcn_perm(Q[3], Q[2], 0x06020400u), y23 = __builtin_amdgcn_perm(Q[3], Q[2], 0x07030501u);
    w[g] = __builtin_amdgcn_perm(x23, x01, 0x05040100u);      // byte column 0
    w[4 + g] = __builtin_amdgcn_perm(x23, x01, 0x07060302u);  // byte column 2
    w[2 + g] = __builtin_amdgcn_perm(y23, y01, 0x05040100u);  // byte column 1
    w[6 + g] = __builtin_amdgcn_perm(y23, y01, 0x07060302u);  // byte column 3
  }
}

// The bit-plane LDE images of the 64-bit rows a and b into ah[16], bh[16]: the
// XOR over the 16 nibble positions P of the table entries (64 bytes each, four
// 16-byte quarters q).  A table address is nibble * 16 plus an immediate: one
// SDWA byte-select AND per nibble (the low nibbles from the row shifted left by
// 4 once).  Unrolled over P; the loads of a position pair are issued together.
__device__ __forceinline__ void lde_planes(uint64_t a, uint64_t b, const uint4* __restrict__ lde16, uint32_t* ah,
                                           uint32_t* bh) {
#pragma unroll
  for (int q = 0; q < 16; ++q) ah[q] = bh[q] = 0u;
#pragma unroll
  for (int h = 0; h < 2; ++h) {  // nibbles 8h .. 8h + 7: word h of the rows
    const uint32_t aw = (uint32_t)(a >> (32 * h)), bw = (uint32_t)(b >> (32 * h));
    const uint32_t aw4 = aw << 4, bw4 = bw << 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // byte j: nibble positions P = 8h + 2j (low), P + 1 (high)
      uint32_t oa0, oa1, ob0, ob1;
      switch (j) {
        case 0: oa0 = hinib_x16<0>(aw4); oa1 = hinib_x16<0>(aw); ob0 = hinib_x16<0>(bw4); ob1 = hinib_x16<0>(bw); break;
        case 1: oa0 = hinib_x16<1>(aw4); oa1 = hinib_x16<1>(aw); ob0 = hinib_x16<1>(bw4); ob1 = hinib_x16<1>(bw); break;
        case 2: oa0 = hinib_x16<2>(aw4); oa1 = hinib_x16<2>(aw); ob0 = hinib_x16<2>(bw4); ob1 = hinib_x16<2>(bw); break;
        default: oa0 = hinib_x16<3>(aw4); oa1 = hinib_x16<3>(aw); ob0 = hinib_x16<3>(bw4); ob1 = hinib_x16<3>(bw); break;
      }
      const int P = 8 * h + 2 * j;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint4* c0 = lde16 + (P * 4 + q) * 16;
        const uint4* c1 = c0 + 64;
        const uint4 x0 = lds_entry(c0, oa0), x1 = lds_entry(c1, oa1), y0 = lds_entry(c0, ob0), y1 = lds_entry(c1, ob1);
        ah[4 * q] = xor3(ah[4 * q], x0.x, x1.x);
        ah[4 * q + 1] = xor3(ah[4 * q + 1], x0.y, x1.y);
        ah[4 * q + 2] = xor3(ah[4 * q + 2], x0.z, x1.z);
        ah[4 * q + 3] = xor3(ah[4 * q + 3], x0.w, x1.w);
        bh[4 * q] = xor3(bh[4 * q], y0.x, y1.x);
        bh[4 * q + 1] = xor3(bh[4 * q + 1], y0.y, y1.y);
        bh[4 * q + 2] = xor3(bh[4 * q + 2], y0.z, y1.z);
        bh[4 * q + 3] = xor3(bh[4 * q + 3], y0.w, y1.w);
      }
    }
  }
}

// Orders the LDS accesses of one wave: the compiler moves no memory access
// across it and the wave's outstanding LDS operations have completed, so data
// one lane stored is visible to every lane of the wave afterwards.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// One wave takes four blocks per pass (quads g = blockIdx.x * R1_NW + wave, then
// strides by gridDim.x * R1_NW); lane (b, J) owns rows (K, J), K < 8, of block
// 4g + b:
//   phase 1, for K = 7 .. 0: the LDE of the row's a and b bits as bit planes
//     (nibble tables, conflict-free), v = a * b on the planes (no lookups), and
//     the Horner step u = alpha u + v, so u = sum_K alpha^K v_K without leaving
//     the lane; then u to bytes, stored to the wave's exchange;
//   phase 2, lane l, per block: W_l = sum_J x^J phi8(u_{J,l}) by two lookups per J
//     in J's table, then acc_l += eq_outer[block] * W_l (comb4p_mul_add).
// out: work-group partials, out[l * gridDim.x + blockIdx.x] for the 64 outputs l.
__global__ void __launch_bounds__(R1_NW * 64) zc_round1_kernel(const uint64_t* __restrict__ A,
                                                              const uint64_t* __restrict__ B, uint64_t n_blocks,
                                                              const uint4* __restrict__ eq_outer,
                                                              const Round1Tables* __restrict__ T,
                                                              uint4* __restrict__ out) {
  constexpr int R1_WG = R1_NW * 64;
  // LDE contributions, chunk-major: lde16[(p * 4 + q) * 16 + v] is words 4q..4q+3
  // of the bit-plane image of nibble value v at nibble position p.  For a fixed
  // (p, q) the 16 values sit in 16 distinct LDS bank groups, so the data-indexed
  // lookups of a wave never conflict (16 KiB).
  struct alignas(16) Lds {
    uint4 lde16[16 * 4 * 16];
    // x^J * phi8(v) = xj[32 J + (v & 15)] ^ xj[32 J + 16 + (v >> 4)] (reduced); J is
    // wave-uniform in phase 2, so each lookup stays in one 16-entry table
    uint4 xj[16 * 32];
    uint4 etab[R1_NW][R1_QUAD][32];  // eq_outer * n and eq_outer * x^4 * n (comb4p_mul_add)
    uint32_t xs[R1_NW][R1_QUAD * 16 * R1_XS];  // u bytes of (b, J) at (16 b + J) * R1_XS
  };
  __shared__ Lds L;
  uint4* const lde16 = L.lde16;
  uint4* const xj = L.xj;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, t = tid & 63u;
  const uint32_t b = t >> 4, J = t & 15u;
  uint4(*const etab)[32] = L.etab[wave];
  uint32_t* const xs = L.xs[wave];
  {
    const uint4* src = reinterpret_cast<const uint4*>(T->lde);  // [p][v][q]
    for (uint32_t i = tid; i < 16 * 4 * 16; i += R1_WG) {
      const uint32_t p = i >> 6, q = (i >> 4) & 3u, v = i & 15u;
      lde16[i] = src[(p * 16 + v) * 4 + q];
    }
    for (uint32_t i = tid; i < 16 * 32; i += R1_WG) xj[i] = T->xj[i];
  }
  uint4 acc = f_zero();
  const uint64_t n_quads = (n_blocks + R1_QUAD - 1) / R1_QUAD;
  __syncthreads();  // the tables are read-only from here; the waves run apart

  for (uint64_t g = (uint64_t)blockIdx.x * R1_NW + wave; g < n_quads; g += (uint64_t)gridDim.x * R1_NW) {
    const uint64_t blk = R1_QUAD * g + b;
    const bool live = blk < n_blocks;  // a quad past the last block (fewer than 4 blocks) adds zero
    {
      const uint4 e = live ? eq_outer[blk] : f_zero();
      etab[b][J] = comb_entry(e, J);
      etab[b][16 + J] = comb_entry(f_mul_x4(e), J);
    }
    // rows 8J + K, K = 7 .. 0, of the block (64 contiguous bytes of each bit
    // vector), each loaded one step ahead
    const uint64_t* const ap = A + blk * R1_ROWS + 8 * J;
    const uint64_t* const bp = B + blk * R1_ROWS + 8 * J;
    uint64_t a_nx = live ? ap[7] : 0ull, b_nx = live ? bp[7] : 0ull;
    uint32_t u[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) u[q] = 0u;
#pragma unroll 1
    for (int K = 7; K >= 0; --K) {
      const uint64_t a = a_nx, bb = b_nx;
      if (live && K > 0) {
        a_nx = ap[K - 1];
        b_nx = bp[K - 1];
      }
      uint32_t ah[16], bh[16];
      lde_planes(a, bb, lde16, ah, bh);
      // planes 8h .. 8h + 7 hold the points l = 32h .. 32h + 31
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        uint32_t v[8];
        gf8_mul_planes(ah + 8 * h, bh + 8 * h, v);
        gf8_horner_planes(u + 8 * h, v);
      }
    }
    uint32_t ub[16];
    planes_to_bytes(u, ub);
    planes_to_bytes(u + 8, ub + 8);
#pragma unroll
    for (int q = 0; q < 16; ++q) xs[(16 * b + J) * R1_XS + q] = ub[q];
    wave_lds_sync();
    // ---- phase 2: lane l ----
    const uint32_t l = t, lw = l >> 2, lsh = 8 * (l & 3u);
#pragma unroll 1
    for (uint32_t b2 = 0; b2 < R1_QUAD; ++b2) {
      uint4 W = f_zero();
#pragma unroll 4
      for (uint32_t j2 = 0; j2 < 16; ++j2) {
        const uint32_t byte = (xs[(16 * b2 + j2) * R1_XS + lw] >> lsh) & 255u;
        const uint4* xt = xj + 32 * j2;
        const uint4 f0 = xt[byte & 15u], f1 = xt[16 + (byte >> 4)];
        W = make_uint4(xor3(W.x, f0.x, f1.x), xor3(W.y, f0.y, f1.y), xor3(W.z, f0.z, f1.z), xor3(W.w, f0.w, f1.w));
      }
      acc = comb4p_mul_add(W, acc, etab[b2]);
    }
    wave_lds_sync();  // the next pass overwrites the exchange and etab
  }
  // Sum the waves' accumulators for output l = t, then store the work-group's
  // partial at out[l * gridDim.x + blockIdx.x].
  __syncthreads();
  reinterpret_cast<uint4*>(L.xs[wave])[t] = acc;
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 1; k < R1_NW; ++k) acc = f_add(acc, reinterpret_cast<const uint4*>(L.xs[k])[t]);
    out[(size_t)t * gridDim.x + blockIdx.x] = acc;
  }
}

// P_C[l] = sum_s phi8(M[s][l]) * Q[s], written to out_c[l].  S holds the
// column sums over row pairs (S[s]: even rows, S[64 + s]: odd rows, each
// weighted by eq over r[7..m)), so Q[s] = S[s] + r6 * (S[s] + S[64 + s]).
// R[s][i] = phi8(2^i) * Q[s] (two products per thread), then thread (l, part)
// XORs the R[s][i] selected by the bits of M[s][l] for its 16 s, and the four
// parts of each l are summed.
__global__ void __launch_bounds__(256) zc_round1_lift_c_kernel(const uint4* __restrict__ S, uint4 r6,
                                                              const Round1Tables* __restrict__ T,
                                                              uint4* __restrict__ out_c) {
  __shared__ uint4 tabs[9 * 16];
  __shared__ uint4 Q[64];
  __shared__ uint4 R[64 * 8];
  __shared__ uint4 part_sum[3][64];
  const uint32_t t = threadIdx.x;
  if (t < 144) tabs[t] = comb_entry(t < 128 ? T->phi_pow[t >> 4] : r6, t & 15u);
  __syncthreads();
  if (t < 64) {
    const uint4 se = S[t], so = S[64 + t];
    Q[t] = f_add(se, comb_mul<1>(f_add(se, so), tabs + 128));
  }
  __syncthreads();
  for (uint32_t e = t; e < 64 * 8; e += 256) R[e] = comb_mul<1>(Q[e >> 3], tabs + (e & 7u) * 16);
  __syncthreads();
  const uint32_t l = t & 63u, part = t >> 6;
  uint4 acc = f_zero();
  for (uint32_t s = 16 * part; s < 16 * part + 16; ++s) {
    const uint32_t m = T->M[s * 64 + l];
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if ((m >> i) & 1u) acc = f_add(acc, R[s * 8 + i]);
  }
  if (part) part_sum[part - 1][l] = acc;
  __syncthreads();
  if (part == 0) out_c[l] = f_add(f_add(acc, part_sum[0][l]), f_add(part_sum[1][l], part_sum[2][l]));
}

__global__ void scale_kernel(uint4* __restrict__ v, uint32_t n, uint4 c) {
  __shared__ uint4 tab[16];
  if (threadIdx.x < 16) tab[threadIdx.x] = comb_entry(c, threadIdx.x);
  __syncthreads();
  if (threadIdx.x < n) v[threadIdx.x] = comb_mul<1>(v[threadIdx.x], tab);
}

}  // namespace flk


namespace flk {
int zc_round1_keep_s(const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                     const flk_f128* r_outer, flk_f128* out_ab, flk_f128* out_c, flk_f128* s_out, void* stream);
}

extern "C" int flk_zc_round1(const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                             const flk_f128* r_outer, flk_f128* out_ab, flk_f128* out_c, void* stream) {
  return flk::zc_round1_keep_s(a_bits, b_bits, c_bits, m, r_outer, out_ab, out_c, nullptr, stream);
}

// flk_zc_round1, and with s_out (device, 128 entries) the column sums S of c
// kept there: S = fold_1b_rows(c as 128-bit elements, eq over r[7..m)).
int flk::zc_round1_keep_s(const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                          const flk_f128* r_outer, flk_f128* out_ab, flk_f128* out_c, flk_f128* s_out,
                          void* stream) {
  using namespace flk;
  int rc = check_enabled();
  if (rc) return rc;
  if (!a_bits || !b_bits || !c_bits || !out_ab || !out_c || (m > 13 && !r_outer))
    return set_error(FLK_E_SHAPE, "zc_round1: null pointer");
  if (m < 13 || m > 40) return set_error(FLK_E_SHAPE, "zc_round1: m=%zu out of range (13..40)", m);
  if (((uintptr_t)a_bits | (uintptr_t)b_bits | (uintptr_t)c_bits) % 16)
    return set_error(FLK_E_SHAPE, "zc_round1: a_bits, b_bits and c_bits must be 16-byte aligned");
  if (out_c != out_ab + 64) return set_error(FLK_E_SHAPE, "zc_round1: out_c must follow out_ab (one 128-entry buffer)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const Round1Tables* T = round1_tables_device(&rc);
  if (rc) return rc;
  // eq over r[7..m) = (fixed r[7..13), r_outer) for the column sums of c, split
  // as eq_lo (2^lo_log) x eq_hi when there are at least 2^8 row pairs.
  flk_f128 fixed[7];
  zc_fixed_challenges(fixed);
  std::vector<flk_f128> r7(fixed + 1, fixed + 7);
  r7.insert(r7.end(), r_outer, r_outer + (m - 13));
  const size_t n_pairs_log = m - 7;
  const bool split = n_pairs_log >= 8;
  const size_t lo_log = split ? (n_pairs_log < 12 ? n_pairs_log : 12) : n_pairs_log;
  const size_t hi_log = n_pairs_log - lo_log;
  const size_t n_outer = size_t(1) << (m - 13);
  const size_t n_lo = size_t(1) << lo_log, n_hi = split ? size_t(1) << hi_log : 0;
  uint4* eq = nullptr;  // [eq_outer | eq_lo | eq_hi | S (128)]
  hipError_t e = flk::stream_alloc(reinterpret_cast<void**>(&eq), (n_outer + n_lo + n_hi + 128) * sizeof(uint4), st);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return set_error(FLK_E_NOMEM, "zc_round1: scratch allocation failed");
  }
  uint4* const eq_lo = eq + n_outer;
  uint4* const eq_hi = eq_lo + n_lo;
  uint4* const S = s_out ? reinterpret_cast<uint4*>(s_out) : eq_hi + n_hi;
  auto fail = [&](int status) {
    (void)flk::stream_free(eq, st);
    return status;
  };
  if ((rc = flk_build_eq(r_outer, m - 13, reinterpret_cast<flk_f128*>(eq), stream))) return fail(rc);
  if ((rc = flk_build_eq(r7.data(), lo_log, reinterpret_cast<flk_f128*>(eq_lo), stream))) return fail(rc);
  const flk_f128* c128 = reinterpret_cast<const flk_f128*>(c_bits);
  if (split) {
    if ((rc = flk_build_eq(r7.data() + lo_log, hi_log, reinterpret_cast<flk_f128*>(eq_hi), stream))) return fail(rc);
    rc = rs_fold_1b_rows_split(c128, reinterpret_cast<const flk_f128*>(eq_lo), lo_log,
                               reinterpret_cast<const flk_f128*>(eq_hi), hi_log, reinterpret_cast<flk_f128*>(S), st);
  } else {
    rc = flk_rs_fold_1b_rows(c128, reinterpret_cast<const flk_f128*>(eq_lo), n_lo, reinterpret_cast<flk_f128*>(S),
                             stream);
  }
  if (rc) return fail(rc);
  uint4* out = reinterpret_cast<uint4*>(out_ab);
  hipLaunchKernelGGL(zc_round1_lift_c_kernel, dim3(1), dim3(256), 0, st, S, u4_of(fixed[0]), T, out + 64);
  FLK_LAUNCH_CHECK("zc_round1_lift_c_kernel");
  // Quads of blocks, one per wave pass; at most 1024 work-groups (two resident
  // per CU, four rounds of them), so every wave takes several quads.
  const uint64_t n_quads = (n_outer + R1_QUAD - 1) / R1_QUAD;
  const uint64_t n_iter = (n_quads + R1_NW - 1) / R1_NW;
  const unsigned grid = (unsigned)(n_iter < 1024 ? n_iter : 1024);
  uint4* parts = nullptr;
  if ((rc = partials_alloc(&parts, grid, 64, out, st))) return fail(rc);
  hipLaunchKernelGGL(zc_round1_kernel, dim3(grid), dim3(R1_NW * 64), 0, st, reinterpret_cast<const uint64_t*>(a_bits),
                     reinterpret_cast<const uint64_t*>(b_bits), (uint64_t)n_outer, eq, T, parts);
  FLK_LAUNCH_CHECK("zc_round1_kernel");
  if ((rc = partials_sum_free(parts, grid, 64, out, st))) return fail(rc);
  const flk_f128 sc = round1_scale();
  hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(64), 0, st, out, 64u, u4_of(sc));
  FLK_LAUNCH_CHECK("scale_kernel");
  e = flk::stream_free(eq, st);
  if (e != hipSuccess) return device_failure(e, "zc_round1: free scratch");
  return FLK_OK;
}
