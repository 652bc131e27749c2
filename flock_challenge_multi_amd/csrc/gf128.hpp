// GF(2^128) arithmetic for gfx950 (GHASH polynomial x^128 + x^7 + x^2 + x + 1,
// polynomial basis, CF/field/gf2_128.rs:10-27).
//
// An element lives in four 32-bit words (uint4: x = bits 0..31 ... w = bits
// 96..127), which is exactly the reference's in-memory layout (lo u64 LE then
// hi u64 LE), so a 16-byte global load is one element.
//
// CDNA4 has no carry-less multiply instruction.  Products are combs: tables of
// multiples of one operand live in LDS, the other operand is cut into nibbles
// or bytes that index them, and the entries are XOR-ed at their offsets into a
// 256-bit accumulator.  The accumulator is congruent to the product mod p, so
// sums of products are XOR-accumulated unreduced and reduced once (the
// deferred-reduction identity of CF/field/gf2_128.rs:46-53).  Forms:
//   comb_mul*      16-entry table k*t, 4-bit Horner (per-lane tables too);
//   comb8_mul*     256-entry byte table (16 lookups; data-indexed 4 KiB tables
//                  collide in the LDS banks);
//   comb4p_mul_add k*t and k*t*x^4, byte Horner (32 conflict-free lookups);
//   combjr_mul_add eight nibble-position tables k*t*x^(4j), no shifts
//                  (32 conflict-free lookups);
//   comb8r_mul_add the byte comb with three-input XORs.
// A 16-entry table of 16-byte entries covers the LDS bank groups once, so its
// data-indexed lookups never conflict.
//
// All routines are plain integer VALU work; nothing here touches MFMA.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flk {

struct U256 {
  uint32_t w[8];
};

__device__ __forceinline__ uint4 f_zero() { return make_uint4(0u, 0u, 0u, 0u); }
__device__ __forceinline__ uint4 f_add(uint4 a, uint4 b) {
  return make_uint4(a.x ^ b.x, a.y ^ b.y, a.z ^ b.z, a.w ^ b.w);
}

// (hi:lo) >> s for 0 < s < 32, i.e. hi << (32-s) | lo >> s.
__device__ __forceinline__ uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t s) {
  return __builtin_amdgcn_alignbit(hi, lo, s);
}

// a ^ b ^ c in one v_bitop3_b32 (truth table 0x96); the compiler does not
// fuse two-input XOR chains into it by itself.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c) {
  return __builtin_amdgcn_bitop3_b32(a, b, c, 0x96);
}

// x * a (CF/field/gf2_128.rs:174-184).
__device__ __forceinline__ uint4 f_mul_x(uint4 a) {
  uint32_t carry = a.w >> 31;
  uint4 o;
  o.w = funnel(a.w, a.z, 31);
  o.z = funnel(a.z, a.y, 31);
  o.y = funnel(a.y, a.x, 31);
  o.x = (a.x << 1) ^ (carry ? 0x87u : 0u);
  return o;
}

__device__ __forceinline__ void u256_zero(U256& a) {
#pragma unroll
  for (int i = 0; i < 8; ++i) a.w[i] = 0u;
}
__device__ __forceinline__ void u256_xor(U256& a, const U256& b) {
#pragma unroll
  for (int i = 0; i < 8; ++i) a.w[i] ^= b.w[i];
}

// Fold bits 128..255 into 0..127 mod p: U*x^128 = U ^ U<<1 ^ U<<2 ^ U<<7, with
// the 7 bits that leave position 127 folded once more
// (same identity as ghash_reduce, CF/field/gf2_128.rs:242-262).
__device__ __forceinline__ uint4 f_reduce(const U256& r) {
  const uint32_t u0 = r.w[4], u1 = r.w[5], u2 = r.w[6], u3 = r.w[7];
  uint4 t;
  t.x = u0 ^ (u0 << 1) ^ (u0 << 2) ^ (u0 << 7);
  t.y = u1 ^ funnel(u1, u0, 31) ^ funnel(u1, u0, 30) ^ funnel(u1, u0, 25);
  t.z = u2 ^ funnel(u2, u1, 31) ^ funnel(u2, u1, 30) ^ funnel(u2, u1, 25);
  t.w = u3 ^ funnel(u3, u2, 31) ^ funnel(u3, u2, 30) ^ funnel(u3, u2, 25);
  const uint32_t ov = (u3 >> 31) ^ (u3 >> 30) ^ (u3 >> 25);
  const uint32_t corr = ov ^ (ov << 1) ^ (ov << 2) ^ (ov << 7);
  return make_uint4(r.w[0] ^ t.x ^ corr, r.w[1] ^ t.y, r.w[2] ^ t.z, r.w[3] ^ t.w);
}

// Entry k (0..15) of the comb table of t: k * t = XOR of t*x^j over set bits j.
__device__ __forceinline__ uint4 comb_entry(uint4 t, uint32_t k) {
  uint4 acc = f_zero();
  uint4 p = t;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if ((k >> j) & 1u) acc = f_add(acc, p);
    if (j < 3) p = f_mul_x(p);
  }
  return acc;
}

// All 16 entries of k * t into registers (used by threads that own a table).
__device__ __forceinline__ void comb_entries(uint4 t, uint4 e[16]) {
  uint4 b0 = t, b1 = f_mul_x(b0), b2 = f_mul_x(b1), b3 = f_mul_x(b2);
  e[0] = f_zero();
  e[1] = b0;
  e[2] = b1;
  e[3] = f_add(b1, b0);
  e[4] = b2;
  e[5] = f_add(b2, b0);
  e[6] = f_add(b2, b1);
  e[7] = f_add(e[6], b0);
  e[8] = b3;
#pragma unroll
  for (int k = 9; k < 16; ++k) e[k] = f_add(b3, e[k - 8]);
}

// Unreduced comb product v * t, where tab[k * stride] = k * t (LDS).
// The result is congruent to v * t mod p; XOR several and call f_reduce once.
template <int STRIDE>
__device__ __forceinline__ void comb_mul_acc(uint4 v, const uint4* __restrict__ tab, U256& acc) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
  U256 a;
  u256_zero(a);
  // Not unrolled over j: keeps only four table loads in flight per lane, which
  // bounds register pressure when several products are inlined together.
#pragma unroll 1
  for (int j = 7; j >= 0; --j) {
    if (j != 7) {
#pragma unroll
      for (int i = 7; i >= 1; --i) a.w[i] = funnel(a.w[i], a.w[i - 1], 28);
      a.w[0] <<= 4;
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t nib = (vw[w] >> (4 * j)) & 15u;
      const uint4 e = tab[nib * STRIDE];
      a.w[w] ^= e.x;
      a.w[w + 1] ^= e.y;
      a.w[w + 2] ^= e.z;
      a.w[w + 3] ^= e.w;
    }
  }
  u256_xor(acc, a);
}

// Same with a run-time stride (per-thread tables laid out entry-major).
// Each Horner step folds the four looked-up entries into the 7-word window
// with three-input XORs.  UNROLL = 1 keeps only four table loads in flight per
// lane, which bounds register pressure when several products are inlined
// together; UNROLL = 8 lets the loads of later steps issue ahead of this step's
// XORs (for kernels whose per-lane tables already hold them at two waves per
// SIMD, where the registers are there).
template <int UNROLL = 1>
__device__ __forceinline__ void comb_mul_acc_rt(uint4 v, const uint4* __restrict__ tab, uint32_t stride,
                                                U256& acc) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
  uint32_t a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 0u;
#pragma unroll UNROLL
  for (int j = 7; j >= 0; --j) {
    // entry nib of the lane's table: one bit-field extract and one shift-add
    const uint4 e0 = tab[__builtin_amdgcn_ubfe(vw[0], 4 * j, 4) * stride];
    const uint4 e1 = tab[__builtin_amdgcn_ubfe(vw[1], 4 * j, 4) * stride];
    const uint4 e2 = tab[__builtin_amdgcn_ubfe(vw[2], 4 * j, 4) * stride];
    const uint4 e3 = tab[__builtin_amdgcn_ubfe(vw[3], 4 * j, 4) * stride];
#pragma unroll
    for (int i = 7; i >= 1; --i) a[i] = funnel(a[i], a[i - 1], 28);
    a[0] = (a[0] << 4) ^ e0.x;
    a[1] = xor3(a[1], e0.y, e1.x);
    a[2] = xor3(a[2], e0.z, e1.y) ^ e2.x;
    a[3] = xor3(xor3(a[3], e0.w, e1.z), e2.y, e3.x);
    a[4] = xor3(a[4], e1.w, e2.z) ^ e3.y;
    a[5] = xor3(a[5], e2.w, e3.z);
    a[6] ^= e3.w;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) acc.w[i] ^= a[i];
}

// ---------------------------------------------------------------------------
// 3-bit comb on per-lane tables of eight entries (tab[k * stride] = k * t,
// k < 8).  Each word of v is cut into eleven windows (bits 3j..3j+2 for j < 10,
// bits 30..31 for j = 10), so a product takes 44 lookups against the 4-bit
// comb's 32, but a table is half the size: seven stores per product (entry 0 is
// a zero row written once, table3_init) and 8 KiB of LDS per wave, so twice the
// waves fit a CU.  The data x data products of the round kernels are bound by
// their LDS stores and latency rather than by VALU issue, and this form runs
// them ~20% faster.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void table3_init(uint4* tab_base, uint32_t tid) { tab_base[tid] = f_zero(); }

// Entries 1..7 of the lane's table of t (entry 0: table3_init).
__device__ __forceinline__ void table3_store_private(uint4* tab_base, uint32_t nthreads, uint32_t tid, uint4 t) {
  const uint4 b0 = t, b1 = f_mul_x(b0), b2 = f_mul_x(b1);
  const uint4 e6 = f_add(b2, b1);
  tab_base[1 * nthreads + tid] = b0;
  tab_base[2 * nthreads + tid] = b1;
  tab_base[3 * nthreads + tid] = f_add(b1, b0);
  tab_base[4 * nthreads + tid] = b2;
  tab_base[5 * nthreads + tid] = f_add(b2, b0);
  tab_base[6 * nthreads + tid] = e6;
  tab_base[7 * nthreads + tid] = f_add(e6, b0);
}

// acc += v * t (unreduced): Horner over the windows from the top, the
// accumulator shifted by 3 between steps.
template <int UNROLL = 1>
__device__ __forceinline__ void comb3_mul_acc_rt(uint4 v, const uint4* __restrict__ tab, uint32_t stride,
                                                 U256& acc) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
  uint32_t a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 0u;
#pragma unroll UNROLL
  for (int j = 10; j >= 0; --j) {
    const uint32_t width = j == 10 ? 2u : 3u;
    const uint4 e0 = tab[__builtin_amdgcn_ubfe(vw[0], 3 * j, width) * stride];
    const uint4 e1 = tab[__builtin_amdgcn_ubfe(vw[1], 3 * j, width) * stride];
    const uint4 e2 = tab[__builtin_amdgcn_ubfe(vw[2], 3 * j, width) * stride];
    const uint4 e3 = tab[__builtin_amdgcn_ubfe(vw[3], 3 * j, width) * stride];
    if (j != 10) {
#pragma unroll
      for (int i = 7; i >= 1; --i) a[i] = funnel(a[i], a[i - 1], 29);
      a[0] <<= 3;
    }
    a[0] ^= e0.x;
    a[1] = xor3(a[1], e0.y, e1.x);
    a[2] = xor3(a[2], e0.z, e1.y) ^ e2.x;
    a[3] = xor3(xor3(a[3], e0.w, e1.z), e2.y, e3.x);
    a[4] = xor3(a[4], e1.w, e2.z) ^ e3.y;
    a[5] = xor3(a[5], e2.w, e3.z);
    a[6] ^= e3.w;
  }
#pragma unroll
  for (int i = 0; i < 8; ++i) acc.w[i] ^= a[i];
}

__device__ __forceinline__ uint4 comb3_mul_rt(uint4 v, const uint4* __restrict__ tab, uint32_t stride) {
  U256 a;
  u256_zero(a);
  comb3_mul_acc_rt(v, tab, stride, a);
  return f_reduce(a);
}

template <int STRIDE>
__device__ __forceinline__ uint4 comb_mul(uint4 v, const uint4* __restrict__ tab) {
  U256 a;
  u256_zero(a);
  comb_mul_acc<STRIDE>(v, tab, a);
  return f_reduce(a);
}

__device__ __forceinline__ uint4 comb_mul_rt(uint4 v, const uint4* __restrict__ tab, uint32_t stride) {
  U256 a;
  u256_zero(a);
  comb_mul_acc_rt(v, tab, stride, a);
  return f_reduce(a);
}

// Unreduced byte-comb product v * t, where tab[b] = b * t for b in 0..255 (LDS,
// 4 KiB).  Sixteen lookups (one per byte of v) and three 8-bit shifts of the
// accumulator, versus 32 lookups and seven shifts for the 4-bit comb.
__device__ __forceinline__ void comb8_mul_acc(uint4 v, const uint4* __restrict__ tab, U256& acc) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
  U256 a;
  u256_zero(a);
#pragma unroll 1
  for (int j = 3; j >= 0; --j) {
    if (j != 3) {
#pragma unroll
      for (int i = 7; i >= 1; --i) a.w[i] = funnel(a.w[i], a.w[i - 1], 24);
      a.w[0] <<= 8;
    }
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const uint32_t byte = (vw[w] >> (8 * j)) & 255u;
      const uint4 e = tab[byte];
      a.w[w] ^= e.x;
      a.w[w + 1] ^= e.y;
      a.w[w + 2] ^= e.z;
      a.w[w + 3] ^= e.w;
    }
  }
  u256_xor(acc, a);
}

__device__ __forceinline__ uint4 comb8_mul(uint4 v, const uint4* __restrict__ tab) {
  U256 a;
  u256_zero(a);
  comb8_mul_acc(v, tab, a);
  return f_reduce(a);
}

// t * x^s for 0 <= s < 32: shift by s, fold the bits above 127 once
// (U x^128 = U ^ U<<1 ^ U<<2 ^ U<<7 for the <= 31-bit overflow U).
__device__ __forceinline__ uint4 f_mul_xs(uint4 t, uint32_t s) {
  if (s == 0) return t;
  const uint32_t o = t.w >> (32 - s);  // bits that leave position 127
  uint4 r;
  r.w = funnel(t.w, t.z, 32 - s);
  r.z = funnel(t.z, t.y, 32 - s);
  r.y = funnel(t.y, t.x, 32 - s);
  r.x = t.x << s;
  // o * (x^7 + x^2 + x + 1): up to 38 bits, words 0 and 1
  r.x ^= o ^ (o << 1) ^ (o << 2) ^ (o << 7);
  r.y ^= (o >> 31) ^ (o >> 30) ^ (o >> 25);
  return r;
}

// t * x^4 (four doublings).
__device__ __forceinline__ uint4 f_mul_x4(uint4 a) { return f_mul_x(f_mul_x(f_mul_x(f_mul_x(a)))); }

// Per-thread private table: entry k of thread `tid` at tab_base[k * nthreads + tid]
// (entry-major, so a wave's 64 lanes hit 64 consecutive 16-byte slots).
__device__ __forceinline__ void table_store_private(uint4* tab_base, uint32_t nthreads, uint32_t tid, uint4 t) {
  uint4 e[16];
  comb_entries(t, e);
#pragma unroll
  for (int k = 0; k < 16; ++k) tab_base[k * nthreads + tid] = e[k];
}

// ---------------------------------------------------------------------------
// gfx950-specific forms.  The three-input XOR V_BITOP3_B32 (truth table 0x96)
// folds two table words per op into the accumulator; the combs below run
// rolled over the four byte positions, taking the top byte of each word per
// step (one SDWA op per table address) and shifting the words up by 8, which
// keeps at most eight table loads in flight and the kernels at <= 128 VGPRs.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 lds_entry(const uint4* tab, uint32_t byte_off) {
  return *reinterpret_cast<const uint4*>(reinterpret_cast<const char*>(tab) + byte_off);
}

// add + (r mod p) for any 256-bit r.  Same folding as f_reduce, with the
// second-order overflow ov (at most 7 bits) merged into u0 before the shifts:
// z = u0 ^ ov, and since ov << 7 stays inside word 0, only u0's own high bits
// cross into word 1 (through the funnels of word 1).
__device__ __forceinline__ uint4 f_reduce_add(const uint32_t r[8], uint4 add) {
  const uint32_t u0 = r[4], u1 = r[5], u2 = r[6], u3 = r[7];
  const uint32_t ov = xor3(u3 >> 31, u3 >> 30, u3 >> 25);
  const uint32_t z = u0 ^ ov;
  uint4 o;
  o.x = xor3(xor3(r[0], add.x, z), z << 1, z << 2) ^ (z << 7);
  o.y = xor3(xor3(r[1], add.y, u1), funnel(u1, u0, 31), funnel(u1, u0, 30)) ^ funnel(u1, u0, 25);
  o.z = xor3(xor3(r[2], add.z, u2), funnel(u2, u1, 31), funnel(u2, u1, 30)) ^ funnel(u2, u1, 25);
  o.w = xor3(xor3(r[3], add.w, u3), funnel(u3, u2, 31), funnel(u3, u2, 30)) ^ funnel(u3, u2, 25);
  return o;
}

// 16 * (high nibble of byte J of w), i.e. (w >> (8J + 4) & 15) << 4.
template <int J>
__device__ __forceinline__ uint32_t hinib_x16(uint32_t w) {
  // One instruction per table address: an SDWA byte select of the AND.  The
  // compiler would fold a preceding shift into the select's shift amount and
  // emit shift + AND instead, so bytes 1-3 are written out.
  uint32_t r;
  if constexpr (J == 1)
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_1 src1_sel:DWORD"
        : "=v"(r) : "v"(w), "v"(0xF0u));
  else if constexpr (J == 2)
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_2 src1_sel:DWORD"
        : "=v"(r) : "v"(w), "v"(0xF0u));
  else if constexpr (J == 3)
    asm("v_and_b32_sdwa %0, %1, %2 dst_sel:DWORD dst_unused:UNUSED_PAD src0_sel:BYTE_3 src1_sel:DWORD"
        : "=v"(r) : "v"(w), "v"(0xF0u));
  else
    r = w & 0xF0u;
  return r;
}

// add + v * t from two 16-entry nibble tables, tab[k] = k * t and
// tab[16 + k] = k * t * x^4: byte J of each word is looked up as its two
// nibbles (both entries at the word's offset) and the bytes are combined by
// Horner in x^8, top byte first.  32 lookups in 512 B of table (built per
// twiddle where a twiddle serves too few products for a 4 KiB byte table).
// The byte loop stays rolled: unrolled, the compiler hoists all 32 loads and
// spills; rolled, at most eight are in flight.
__device__ __forceinline__ uint4 comb4p_mul_add(uint4 v, uint4 add, const uint4* tab) {
  uint32_t hw[4] = {v.x, v.y, v.z, v.w};
  uint32_t lw[4] = {v.x << 4, v.y << 4, v.z << 4, v.w << 4};
  uint32_t a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 0u;
#pragma unroll 1
  for (int J = 0; J < 4; ++J) {
#pragma unroll
    for (int i = 7; i >= 1; --i) a[i] = funnel(a[i], a[i - 1], 24);
    a[0] <<= 8;
    // Low nibbles, then high nibbles: half the entries live at a time.
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      uint4 e[4];
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        e[w] = half ? lds_entry(tab + 16, hinib_x16<3>(hw[w])) : lds_entry(tab, hinib_x16<3>(lw[w]));
        if (half) hw[w] <<= 8; else lw[w] <<= 8;
      }
      a[0] ^= e[0].x;
      a[1] = xor3(a[1], e[0].y, e[1].x);
      a[2] = xor3(a[2], e[0].z, e[1].y) ^ e[2].x;
      a[3] = xor3(xor3(a[3], e[0].w, e[1].z), e[2].y, e[3].x);
      a[4] = xor3(a[4], e[1].w, e[2].z) ^ e[3].y;
      a[5] = xor3(a[5], e[2].w, e[3].z);
      a[6] ^= e[3].w;
    }
  }
  return f_reduce_add(a, add);
}

// Work-group build of the nibble-position tables of n constants t[c] in LDS
// (the combjr_mul_add layout: tab[c * 128 + j * 16 + k] = k * t[c] * x^(4j)),
// through nib[c * 32 + s] = t[c] * x^s.  get(c) returns t[c].  Contains two
// barriers; every thread of the work-group must call it.
template <typename Get>
__device__ __forceinline__ void build_pos_tables(uint4* tab, uint4* nib, uint32_t n, Get get) {
  for (uint32_t e = threadIdx.x; e < n * 32u; e += blockDim.x) nib[e] = f_mul_xs(get(e >> 5), e & 31u);
  __syncthreads();
  for (uint32_t e = threadIdx.x; e < n * 128u; e += blockDim.x) {
    const uint32_t c = e >> 7, j = (e >> 4) & 7u, k = e & 15u;
    uint4 v = f_zero();
#pragma unroll
    for (int b = 0; b < 4; ++b)
      if ((k >> b) & 1u) v = f_add(v, nib[c * 32 + 4 * j + b]);
    tab[e] = v;
  }
  __syncthreads();
}

// acc += v * t unreduced (seven words, bits 0..223), with t's nibble-position
// tables (tab[j * 16 + k] = k * t * x^(4j)): word w of v contributes its eight
// entries at word offset w, so sums of products need no shifts and one
// reduction at the end (f_reduce of the seven words and a zero eighth).
__device__ __forceinline__ void combjr_acc(uint4 v, const uint4* tab, uint32_t acc[7]) {
  const uint32_t vw[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint32_t hi = vw[w], lo = hi << 4;
    const uint4 e0 = lds_entry(tab + 0 * 16, hinib_x16<0>(lo));
    const uint4 e1 = lds_entry(tab + 1 * 16, hinib_x16<0>(hi));
    const uint4 e2 = lds_entry(tab + 2 * 16, hinib_x16<1>(lo));
    const uint4 e3 = lds_entry(tab + 3 * 16, hinib_x16<1>(hi));
    const uint4 e4 = lds_entry(tab + 4 * 16, hinib_x16<2>(lo));
    const uint4 e5 = lds_entry(tab + 5 * 16, hinib_x16<2>(hi));
    const uint4 e6 = lds_entry(tab + 6 * 16, hinib_x16<3>(lo));
    const uint4 e7 = lds_entry(tab + 7 * 16, hinib_x16<3>(hi));
    acc[w + 0] = xor3(xor3(xor3(xor3(acc[w + 0], e0.x, e1.x), e2.x, e3.x), e4.x, e5.x), e6.x, e7.x);
    acc[w + 1] = xor3(xor3(xor3(xor3(acc[w + 1], e0.y, e1.y), e2.y, e3.y), e4.y, e5.y), e6.y, e7.y);
    acc[w + 2] = xor3(xor3(xor3(xor3(acc[w + 2], e0.z, e1.z), e2.z, e3.z), e4.z, e5.z), e6.z, e7.z);
    acc[w + 3] = xor3(xor3(xor3(xor3(acc[w + 3], e0.w, e1.w), e2.w, e3.w), e4.w, e5.w), e6.w, e7.w);
  }
}

// The 32 table offsets of v for the nibble-position combs, top word first:
// ad[8 * h + j] = 16 * (nibble j of word 3 - h of v).
__device__ __forceinline__ void pos_offsets(uint4 v, uint32_t ad[32]) {
  const uint32_t vw[4] = {v.w, v.z, v.y, v.x};
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    const uint32_t hi = vw[h], lo = hi << 4;
    ad[8 * h + 0] = hinib_x16<0>(lo);
    ad[8 * h + 1] = hinib_x16<0>(hi);
    ad[8 * h + 2] = hinib_x16<1>(lo);
    ad[8 * h + 3] = hinib_x16<1>(hi);
    ad[8 * h + 4] = hinib_x16<2>(lo);
    ad[8 * h + 5] = hinib_x16<2>(hi);
    ad[8 * h + 6] = hinib_x16<3>(lo);
    ad[8 * h + 7] = hinib_x16<3>(hi);
  }
}

// v * t from v's precomputed offsets (pos_offsets) and t's nibble-position
// tables (tab[j * 16 + k] = k * t * x^(4j)): the combjr_mul_add product for
// several t with one v.
__device__ __forceinline__ uint4 combjr_mul_pre(const uint32_t ad[32], const uint4* tab) {
  uint32_t a[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) a[i] = 0u;
#pragma unroll
  for (int h = 0; h < 4; ++h) {
#pragma unroll
    for (int i = 6; i >= 1; --i) a[i] = a[i - 1];
    const uint4 e0 = lds_entry(tab + 0 * 16, ad[8 * h + 0]);
    const uint4 e1 = lds_entry(tab + 1 * 16, ad[8 * h + 1]);
    const uint4 e2 = lds_entry(tab + 2 * 16, ad[8 * h + 2]);
    const uint4 e3 = lds_entry(tab + 3 * 16, ad[8 * h + 3]);
    const uint4 e4 = lds_entry(tab + 4 * 16, ad[8 * h + 4]);
    const uint4 e5 = lds_entry(tab + 5 * 16, ad[8 * h + 5]);
    const uint4 e6 = lds_entry(tab + 6 * 16, ad[8 * h + 6]);
    const uint4 e7 = lds_entry(tab + 7 * 16, ad[8 * h + 7]);
    a[0] = xor3(xor3(xor3(e0.x, e1.x, e2.x), e3.x, e4.x), e5.x, e6.x) ^ e7.x;
    a[1] = xor3(xor3(xor3(xor3(a[1], e0.y, e1.y), e2.y, e3.y), e4.y, e5.y), e6.y, e7.y);
    a[2] = xor3(xor3(xor3(xor3(a[2], e0.z, e1.z), e2.z, e3.z), e4.z, e5.z), e6.z, e7.z);
    a[3] = xor3(xor3(xor3(xor3(a[3], e0.w, e1.w), e2.w, e3.w), e4.w, e5.w), e6.w, e7.w);
  }
  const uint32_t r[8] = {a[0], a[1], a[2], a[3], a[4], a[5], a[6], 0u};
  return f_reduce_add(r, f_zero());
}

// XOR of tab[j * 16 + nibble j of w] over the eight nibbles j of w: one word of
// a GF(2)-linear map applied with 16-entry subset-sum tables (conflict-free
// lookups, one address instruction each).
__device__ __forceinline__ uint4 nibtab8_sum(uint32_t w, const uint4* tab) {
  const uint32_t lo = w << 4;
  const uint4 e0 = lds_entry(tab + 0 * 16, hinib_x16<0>(lo));
  const uint4 e1 = lds_entry(tab + 1 * 16, hinib_x16<0>(w));
  const uint4 e2 = lds_entry(tab + 2 * 16, hinib_x16<1>(lo));
  const uint4 e3 = lds_entry(tab + 3 * 16, hinib_x16<1>(w));
  const uint4 e4 = lds_entry(tab + 4 * 16, hinib_x16<2>(lo));
  const uint4 e5 = lds_entry(tab + 5 * 16, hinib_x16<2>(w));
  const uint4 e6 = lds_entry(tab + 6 * 16, hinib_x16<3>(lo));
  const uint4 e7 = lds_entry(tab + 7 * 16, hinib_x16<3>(w));
  uint4 r;
  r.x = xor3(xor3(xor3(e0.x, e1.x, e2.x), e3.x, e4.x), e5.x, e6.x) ^ e7.x;
  r.y = xor3(xor3(xor3(e0.y, e1.y, e2.y), e3.y, e4.y), e5.y, e6.y) ^ e7.y;
  r.z = xor3(xor3(xor3(e0.z, e1.z, e2.z), e3.z, e4.z), e5.z, e6.z) ^ e7.z;
  r.w = xor3(xor3(xor3(e0.w, e1.w, e2.w), e3.w, e4.w), e5.w, e6.w) ^ e7.w;
  return r;
}

// add + v * t with eight nibble-position tables tab[j * 16 + k] = k * t * x^(4j)
// (j = nibble position inside a 32-bit word; 2 KiB per twiddle).  Every
// 16-entry slice covers all LDS bank groups once, so the 32 lookups of a
// product are free of bank conflicts (lanes that hit one entry are served by
// broadcast).  Words are combined by Horner in x^32, top word first (a shift by
// one word is a register move), two words per rolled iteration.
// JS: entries between the starts of consecutive nibble-position tables (16, or
// 17 where the tables are written by lanes that each own one position: the pad
// spreads those writes over all LDS bank groups; lookups only change offsets).
// Rolled over the two word pairs: unrolled (word shifts as register renames
// instead of moves) the NTT's 3-layer pass needs 110 VGPRs, runs at 4 instead of
// 8 waves per SIMD, and measured no faster (3.59 against 3.52 ms per m = 32 encode).
template <int JS = 16>
__device__ __forceinline__ uint4 combjr_mul_add(uint4 v, uint4 add, const uint4* tab) {
  uint32_t vw[4] = {v.w, v.z, v.y, v.x};
  uint32_t a[7];
#pragma unroll
  for (int i = 0; i < 7; ++i) a[i] = 0u;
#pragma unroll 1
  for (int it = 0; it < 2; ++it) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int i = 6; i >= 1; --i) a[i] = a[i - 1];
      const uint32_t hi = vw[h], lo = hi << 4;
      const uint4 e0 = lds_entry(tab + 0 * JS, hinib_x16<0>(lo));
      const uint4 e1 = lds_entry(tab + 1 * JS, hinib_x16<0>(hi));
      const uint4 e2 = lds_entry(tab + 2 * JS, hinib_x16<1>(lo));
      const uint4 e3 = lds_entry(tab + 3 * JS, hinib_x16<1>(hi));
      const uint4 e4 = lds_entry(tab + 4 * JS, hinib_x16<2>(lo));
      const uint4 e5 = lds_entry(tab + 5 * JS, hinib_x16<2>(hi));
      const uint4 e6 = lds_entry(tab + 6 * JS, hinib_x16<3>(lo));
      const uint4 e7 = lds_entry(tab + 7 * JS, hinib_x16<3>(hi));
      a[0] = xor3(xor3(xor3(e0.x, e1.x, e2.x), e3.x, e4.x), e5.x, e6.x) ^ e7.x;
      a[1] = xor3(xor3(xor3(xor3(a[1], e0.y, e1.y), e2.y, e3.y), e4.y, e5.y), e6.y, e7.y);
      a[2] = xor3(xor3(xor3(xor3(a[2], e0.z, e1.z), e2.z, e3.z), e4.z, e5.z), e6.z, e7.z);
      a[3] = xor3(xor3(xor3(xor3(a[3], e0.w, e1.w), e2.w, e3.w), e4.w, e5.w), e6.w, e7.w);
    }
    vw[0] = vw[2];
    vw[1] = vw[3];
  }
  // a holds words 0..6; reduce with word 7 = 0
  const uint32_t r[8] = {a[0], a[1], a[2], a[3], a[4], a[5], a[6], 0u};
  return f_reduce_add(r, add);
}

// add + v * t with tab[b] = b * t (b in 0..255, LDS): the byte comb with
// three-input XORs, rolled over the four byte positions like comb4p_mul_add.
__device__ __forceinline__ uint4 comb8r_mul_add(uint4 v, uint4 add, const uint4* tab) {
  uint32_t vw[4] = {v.x, v.y, v.z, v.w};
  uint32_t a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 0u;
#pragma unroll 1
  for (int J = 0; J < 4; ++J) {
    uint4 e[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      e[w] = lds_entry(tab, (vw[w] >> 20) & 0xFF0u);
      vw[w] <<= 8;
    }
#pragma unroll
    for (int i = 7; i >= 1; --i) a[i] = funnel(a[i], a[i - 1], 24);
    a[0] = (a[0] << 8) ^ e[0].x;
    a[1] = xor3(a[1], e[0].y, e[1].x);
    a[2] = xor3(a[2], e[0].z, e[1].y) ^ e[2].x;
    a[3] = xor3(xor3(a[3], e[0].w, e[1].z), e[2].y, e[3].x);
    a[4] = xor3(a[4], e[1].w, e[2].z) ^ e[3].y;
    a[5] = xor3(a[5], e[2].w, e[3].z);
    a[6] ^= e[3].w;
  }
  return f_reduce_add(a, add);
}

}  // namespace flk
