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
// P_C needs no per-row LDE: the LDE is GF(2^8)-linear and c is a bit, so
//     P_C[l] = sum_s phi8(M[s][l]) * Q[s],   Q[s] = sum_x eq(x) c_x[s],
// with M the 64 x 64 LDE matrix (S -> Lambda).  Q is a column sum of the bit
// matrix c: the ring-switch fold_1b_rows kernel computes it over row pairs (c
// read as 2^(m-7) elements of 128 bits, weighted by eq over r[7..m)), and a
// 64-thread epilogue binds r[6] and applies M.
//
// In the P_AB kernel one wave processes one 128-row block (one x_outer) at a
// time; per block:
//   phase 1, per 64-row half, lane t = row 64 * half + t (row r = (K, J) =
//     (r & 7, r >> 3)): the 64-point LDE of the row's a and b bits (a fixed GF(2)-linear map 64 bits -> 64 bytes, applied
//     with 16 nibble tables of 64 bytes), v_ab = a_l * b_l in GF(2^8) from the
//     full 64 KiB product table (one lookup of a * 256 + b, the index formed by
//     one byte permute; FLOCK_R1_FULLTAB=0 selects split-nibble tables, two
//     lookups), stored to LDS;
//   phase 2a, lane = (byte word w, J-group): Horner over K for the half's 2 J
//     values of its group and 4 l at once (packed bytes); after both halves,
//     x^J phi8 of each byte by two lookups in J's table of x^J phi8(n) and
//     x^J phi8(n << 4) (reduced), summed over its 4 J;
//   phase 2b, lane = l: the sum over the four J-groups, then one product by
//     eq_outer with nibble-pair tables (comb4p_mul_add).
// Every data-indexed 16-byte LDS table has 16 entries (conflict-free); the byte
// product table takes one lookup per product where the split tables take two
// and 2.3x the address arithmetic, which outweighs its bank conflicts.
#include <hip/hip_runtime.h>

#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

struct Round1Tables {
  uint32_t lde[16 * 16 * 16];  // [nibble pos p][value v][word w]: bytes l = 4w..4w+3
  uint8_t mul_lo[256 * 16];    // [a][n] = a * n        in GF(2^8)
  uint8_t mul_hi[256 * 16];    // [a][n] = a * (n << 4) in GF(2^8)
  uint8_t mul64[256 * 256];    // [a][b] = a * b in GF(2^8)
  uint8_t M[64 * 64];          // [s][l]: the LDE of the unit vector e_s at point l
  uint4 xj[16 * 32];           // [J][n] = x^J * phi8(n), [J][16 + n] = x^J * phi8(n << 4), reduced
  uint4 phi_pow[8];            // phi8(2^i)
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
      memcpy(&T.lde[(p * 16 + val) * 16], bytes, 64);
    }
  for (int a = 0; a < 256; ++a)
    for (int n = 0; n < 16; ++n) {
      T.mul_lo[a * 16 + n] = g8mul((uint8_t)a, (uint8_t)n);
      T.mul_hi[a * 16 + n] = g8mul((uint8_t)a, (uint8_t)(n << 4));
    }
  for (int a = 0; a < 256; ++a)
    for (int b = 0; b < 256; ++b) T.mul64[a * 256 + b] = g8mul((uint8_t)a, (uint8_t)b);
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

constexpr int R1_ROWS = 128;  // rows of one block (one x_outer)
// Waves per work-group: each wave processes whole blocks on its own (no
// work-group barrier inside the block loop), so the waves of a CU drift apart
// and one wave's table lookups overlap another's arithmetic.  A wave's block
// scratch holds one 64-row half of a block at a time (5.5 KiB), so with the
// full 64 KiB product table one work-group of 12 waves fits the LDS (146 KiB);
// with the split-nibble tables (8 KiB) two work-groups of 8 waves (69 KiB each).
template <bool FULL>
constexpr int r1_waves() { return FULL ? 12 : 8; }

__device__ __forceinline__ uint32_t xtime4(uint32_t w) {
  return ((w & 0x7F7F7F7Fu) << 1) ^ (((w >> 7) & 0x01010101u) * 0x1Bu);
}

// Orders the LDS accesses of one wave: the compiler moves no memory access
// across it and the wave's outstanding LDS operations have completed, so data
// one lane stored is visible to every lane of the wave afterwards.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Per-wave block scratch: row r (0..63) of the current half at r * 17 (16
// product words; 17 * 16 = 16 mod 64, so the phase-2a column reads of the four
// J-groups fall in distinct banks), later the phase-2a partial sums at
// ypart[(q * 4 + j) * 64 + (lane ^ 16 j)] (the XOR keeps the phase-2b gathers of
// the four j in distinct banks).
constexpr uint32_t R1_RP = 17, R1_YP = 64;
constexpr uint32_t R1_SCRATCH = 64 * R1_RP;  // words; >= 16 * R1_YP

// Each wave takes blocks blockIdx.x * NW + wave, then strides by gridDim.x * NW;
// each block follows the phases described at the top of the file.
// out: work-group partials, out[l * gridDim.x + blockIdx.x] for the 64 outputs l.
template <bool FULL>
__global__ void __launch_bounds__(r1_waves<FULL>() * 64) zc_round1_kernel(
    const uint64_t* __restrict__ A, const uint64_t* __restrict__ B, uint64_t n_blocks,
    const uint4* __restrict__ eq_outer, const Round1Tables* __restrict__ T, uint4* __restrict__ out) {
  constexpr int NW = r1_waves<FULL>(), R1_WG = NW * 64;
  static_assert(64 * R1_RP <= R1_SCRATCH, "a half block's rows must fit in the scratch");
  static_assert(16 * R1_YP <= R1_SCRATCH, "a wave's sums must fit in its scratch");
  // LDE contributions, chunk-major: lde16[(p * 4 + q) * 16 + v] is bytes 16q..16q+15
  // of the 64-byte image of nibble value v at nibble position p.  For a fixed
  // (p, q) the 16 values sit in 16 distinct LDS bank groups, so the data-indexed
  // lookups of a wave never conflict (16 KiB).
  // One LDS block with the byte product tables first: their lookups then fit the
  // 16-bit offset field of ds_read_u8 (no address add per lookup).
  struct alignas(16) Lds {
    // FULL: mul[a * 256 + b] = a * b (one lookup per product); else
    // mul[a * 16 + n] = a * n and mul[4096 + a * 16 + n] = a * (n << 4)
    uint8_t mul[FULL ? 65536 : 8192];
    uint4 lde16[16 * 4 * 16];
    // x^J * phi8(v) = xj[32 J + (v & 15)] ^ xj[32 J + 16 + (v >> 4)] (reduced): the
    // 16 lanes of a quarter-wave share J (one J-group), so their lookups stay in
    // one 16-entry table each
    uint4 xj[16 * 32];
    uint4 etab[NW][32];  // eq_outer * n and eq_outer * x^4 * n (comb4p_mul_add)
    uint32_t scratch[NW][R1_SCRATCH];
  };
  __shared__ Lds L;
  uint4* const lde16 = L.lde16;
  uint8_t* const mul_lo = L.mul;
  uint8_t* const mul_hi = L.mul + 4096;
  uint4* const xj = L.xj;
  const uint32_t tid = threadIdx.x, wave = tid >> 6, t = tid & 63u;
  uint32_t* const rows = L.scratch[wave];
  uint32_t* const ypart = rows;
  uint4* const etab = L.etab[wave];
  {
    const uint4* src = reinterpret_cast<const uint4*>(T->lde);  // [p][v][q]
    for (uint32_t i = tid; i < 16 * 4 * 16; i += R1_WG) {
      const uint32_t p = i >> 6, q = (i >> 4) & 3u, v = i & 15u;
      lde16[i] = src[(p * 16 + v) * 4 + q];
    }
    if (FULL) {
      const uint4* m64 = reinterpret_cast<const uint4*>(T->mul64);
      for (uint32_t i = tid; i < 4096; i += R1_WG) reinterpret_cast<uint4*>(L.mul)[i] = m64[i];
    } else {
      const uint32_t* mlo = reinterpret_cast<const uint32_t*>(T->mul_lo);
      const uint32_t* mhi = reinterpret_cast<const uint32_t*>(T->mul_hi);
      for (uint32_t i = tid; i < 1024; i += R1_WG) {
        reinterpret_cast<uint32_t*>(mul_lo)[i] = mlo[i];
        reinterpret_cast<uint32_t*>(mul_hi)[i] = mhi[i];
      }
    }
    for (uint32_t i = tid; i < 16 * 32; i += R1_WG) xj[i] = T->xj[i];
  }
  // phase-2a role: 16 words x 4 J-groups; the group jg takes J = 2jg, 2jg + 1 of
  // each half (J = 8 * half + 2 * jg + jj)
  const uint32_t w = t & 15u, jg = t >> 4;
  uint4 acc = f_zero();
  __syncthreads();  // the tables are read-only from here; the waves run apart

  for (uint64_t blk = (uint64_t)blockIdx.x * NW + wave; blk < n_blocks; blk += (uint64_t)gridDim.x * NW) {
    // eq_outer[blk] nibble tables for comb4p_mul_add
    if (t < 32) {
      uint4 e = eq_outer[blk];
      if (t >= 16) e = f_mul_x4(e);
      etab[t] = comb_entry(e, t & 15u);
    }
    uint32_t wj[4];
#pragma unroll 1
    for (uint32_t half = 0; half < 2; ++half) {
      // ---- phase 1: row t of the half, block row r = 64 * half + t = (K, J) =
      // (r & 7, r >> 3) ----
      const uint64_t row = blk * R1_ROWS + 64u * half + t;
      const uint64_t a = A[row], b = B[row];
      uint32_t ah[16], bh[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) ah[q] = bh[q] = 0u;
      // Rolled over nibble-position pairs: at most 16 table loads in flight.
#pragma unroll 1
      for (uint32_t p = 0; p < 16; p += 2) {
        const uint32_t na0 = (uint32_t)(a >> (4 * p)) & 15u, na1 = (uint32_t)(a >> (4 * p + 4)) & 15u;
        const uint32_t nb0 = (uint32_t)(b >> (4 * p)) & 15u, nb1 = (uint32_t)(b >> (4 * p + 4)) & 15u;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint4* c0 = lde16 + (p * 4 + q) * 16;
          const uint4* c1 = c0 + 64;
          const uint4 x0 = c0[na0], x1 = c1[na1], y0 = c0[nb0], y1 = c1[nb1];
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
      uint32_t* rd = rows + t * R1_RP;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (FULL) {
            // index (byte j of ah) * 256 + (byte j of bh) in one byte permute
            const uint32_t idx = __builtin_amdgcn_perm(ah[q], bh[q], 0x0c0c0000u | ((4u + j) << 8) | (uint32_t)j);
            v |= (uint32_t)L.mul[idx] << (8 * j);
          } else {
            const uint32_t x16 = ((ah[q] >> (8 * j)) & 255u) << 4, y = (bh[q] >> (8 * j)) & 255u;
            v |= (uint32_t)(mul_lo[x16 | (y & 15u)] ^ mul_hi[x16 | (y >> 4)]) << (8 * j);
          }
        }
        rd[q] = v;
      }
      wave_lds_sync();
      // ---- phase 2a (this half): lane (w, jg): Horner over K in alpha = 0x02
      // (xtime on packed bytes: 4 l per word) for J = 8 * half + 2 * jg + jj ----
      uint32_t h[2];
#pragma unroll
      for (int jj = 0; jj < 2; ++jj) {
        const uint32_t* col = rows + (8u * (2u * jg + jj)) * R1_RP + w;
        h[jj] = 0;
#pragma unroll
        for (int K = 7; K >= 0; --K) h[jj] = xtime4(h[jj]) ^ col[R1_RP * K];
      }
      // constant indices on both sides of the (wave-uniform) branch keep wj in registers
      if (half == 0) {
        wj[0] = h[0];
        wj[1] = h[1];
      } else {
        wj[2] = h[0];
        wj[3] = h[1];
      }
      wave_lds_sync();  // every lane has read the half's rows; the next writes reuse them
    }
    // Y[j] = sum over the lane's four J of x^J phi8(byte j of wj[i]), reduced
    uint4 Y[4] = {f_zero(), f_zero(), f_zero(), f_zero()};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const uint4* xt = xj + 32 * (8 * (i >> 1) + 2 * jg + (i & 1));
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t byte = (wj[i] >> (8 * j)) & 255u;
        const uint4 f0 = xt[byte & 15u], f1 = xt[16 + (byte >> 4)];
        Y[j] = make_uint4(xor3(Y[j].x, f0.x, f1.x), xor3(Y[j].y, f0.y, f1.y), xor3(Y[j].z, f0.z, f1.z),
                          xor3(Y[j].w, f0.w, f1.w));
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ypart[(0 * 4 + j) * R1_YP + (t ^ (16u * j))] = Y[j].x;
      ypart[(1 * 4 + j) * R1_YP + (t ^ (16u * j))] = Y[j].y;
      ypart[(2 * 4 + j) * R1_YP + (t ^ (16u * j))] = Y[j].z;
      ypart[(3 * 4 + j) * R1_YP + (t ^ (16u * j))] = Y[j].w;
    }
    wave_lds_sync();
    // ---- phase 2b: lane l: finish the J sum over the four J-groups, multiply by
    // eq_outer ----
    {
      const uint32_t l = t, ww = l >> 2, j = l & 3u;
      uint32_t y[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        // writer lane of J-group g: ww + 16 g, stored at (ww + 16 g) ^ 16 j
        const uint32_t* yp = ypart + (q * 4 + j) * R1_YP;
        y[q] = xor3(yp[ww ^ (16u * j)], yp[(ww + 16u) ^ (16u * j)], yp[(ww + 32u) ^ (16u * j)]) ^
               yp[(ww + 48u) ^ (16u * j)];
      }
      acc = comb4p_mul_add(make_uint4(y[0], y[1], y[2], y[3]), acc, etab);
    }
    wave_lds_sync();  // the next block overwrites the scratch and etab
  }
  // Sum the waves' accumulators for output l = t, then store the work-group's
  // partial at out[l * gridDim.x + blockIdx.x].
  __syncthreads();
  reinterpret_cast<uint4*>(rows)[t] = acc;
  __syncthreads();
  if (wave == 0) {
#pragma unroll
    for (int k = 1; k < NW; ++k) acc = f_add(acc, reinterpret_cast<const uint4*>(L.scratch[k])[t]);
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
  if (((uintptr_t)a_bits | (uintptr_t)b_bits) % 8 || (uintptr_t)c_bits % 16)
    return set_error(FLK_E_SHAPE, "zc_round1: a_bits and b_bits must be 8-byte aligned, c_bits 16-byte aligned");
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
  // The full 64 KiB product table (one work-group of 12 waves per CU) unless
  // FLOCK_R1_FULLTAB=0 selects the split-nibble tables (two work-groups of 8
  // waves per CU).  Grids of 512 (full) or 1024 (split) work-groups: several
  // blocks per wave, so the waves' uneven block counts even out.
  static const bool full = [] {
    const char* v = getenv("FLOCK_R1_FULLTAB");
    return !(v && v[0] == '0');
  }();
  const int nw = full ? r1_waves<true>() : r1_waves<false>();
  const uint64_t n_iter = (n_outer + nw - 1) / nw;
  const uint64_t cap = full ? 512 : 1024;
  const unsigned grid = (unsigned)(n_iter < cap ? n_iter : cap);
  uint4* parts = nullptr;
  if ((rc = partials_alloc(&parts, grid, 64, out, st))) return fail(rc);
  if (full)
    hipLaunchKernelGGL(zc_round1_kernel<true>, dim3(grid), dim3(nw * 64), 0, st,
                       reinterpret_cast<const uint64_t*>(a_bits), reinterpret_cast<const uint64_t*>(b_bits),
                       (uint64_t)n_outer, eq, T, parts);
  else
    hipLaunchKernelGGL(zc_round1_kernel<false>, dim3(grid), dim3(nw * 64), 0, st,
                       reinterpret_cast<const uint64_t*>(a_bits), reinterpret_cast<const uint64_t*>(b_bits),
                       (uint64_t)n_outer, eq, T, parts);
  FLK_LAUNCH_CHECK("zc_round1_kernel");
  if ((rc = partials_sum_free(parts, grid, 64, out, st))) return fail(rc);
  const flk_f128 sc = round1_scale();
  hipLaunchKernelGGL(scale_kernel, dim3(1), dim3(64), 0, st, out, 64u, u4_of(sc));
  FLK_LAUNCH_CHECK("scale_kernel");
  e = flk::stream_free(eq, st);
  if (e != hipSuccess) return device_failure(e, "zc_round1: free scratch");
  return FLK_OK;
}
