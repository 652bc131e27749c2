// Lincheck, ring-switch and zerocheck evaluation reductions for gfx950.
//
//   flk_lc_partial_fold     partial_fold_packed_z(_fast)  CF/lincheck.rs:2233-2420
//   flk_lc_bind_top         sumcheck_bind_top_in_place    CF/lincheck.rs:4968-4990
//   flk_lc_round_msg_msb    product-sumcheck round eval   CF/lincheck.rs:5010-5030
//   flk_rs_fold_1b_rows     fold_1b_rows                  CF/pcs/ring_switch.rs:1594-1642
//   flk_rs_fold_b128_elems  fold_b128_elems               CF/pcs/ring_switch.rs:1952-2087
//   flk_zc_fold_at_z        fold_at_z                     CF/zerocheck/multilinear.rs:575-597
//   flk_zc_round_pair       round_pair (wire form)        CF/zerocheck/multilinear.rs:616-640,
//                                                         CF/zerocheck.rs:1748-1781
//   flk_zc_fold_round       fold_in_place_pair + round_pair against a split eq
//                                                         CF/zerocheck/multilinear.rs:616-640,3194-3213
//   flk_zc_fold_at_z_round  fold_at_z + round_pair from the bits (with or without the table)
//   flk_zc_fold_at_z_bind_round  fold_at_z + fold_in_place_pair + round_pair from the bits
//                                                         CF/zerocheck/multilinear.rs:575-683,744-790
//
// These are XOR reductions of F2-linear maps over bit-packed witness data.  The
// sums of known values selected by bits ("sum eq[i] over the set bits") use
// subset-sum tables in LDS, the GPU form of the reference's method-of-four-
// Russians tables: a 16-entry table per nibble of selector bits (conflict-free,
// 256 B; byte tables collide in the LDS banks when data index them).  Products
// of two data values use the per-lane comb of gf128.hpp.  Sums over all items
// end as per-work-group partials added by a second launch (partials_sum_free).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>

#include "gf128.hpp"
#include "internal.hpp"

namespace flk {

__device__ __forceinline__ uint4 shfl_xor4_r(uint4 v, int m) {
  return make_uint4(__shfl_xor(v.x, m), __shfl_xor(v.y, m), __shfl_xor(v.z, m), __shfl_xor(v.w, m));
}

__device__ __forceinline__ void atomic_xor4(uint4* out, uint4 v) {
  uint32_t* o = reinterpret_cast<uint32_t*>(out);
  if (v.x) atomicXor(o + 0, v.x);
  if (v.y) atomicXor(o + 1, v.y);
  if (v.z) atomicXor(o + 2, v.z);
  if (v.w) atomicXor(o + 3, v.w);
}

// XOR-reduce over the work-group (blockDim multiple of 64) into *out.
// XOR-reduce v over the work-group; thread 0 stores it at parts[k * gridDim.x +
// blockIdx.x] (the work-group partials that partials_sum_free adds up).
__device__ __forceinline__ void wg_reduce_store(uint4 v, uint4* red, uint4* parts, uint32_t k) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v = f_add(v, shfl_xor4_r(v, m));
  if ((threadIdx.x & 63u) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint4 acc = red[0];
    for (uint32_t w = 1; w < blockDim.x / 64u; ++w) acc = f_add(acc, red[w]);
    parts[(size_t)k * gridDim.x + blockIdx.x] = acc;
  }
}



// ---------------------------------------------------------------------------
// Lincheck partial fold.  out[i] = sum over stripes s and bits r of byte
// z[s*k + i] of eq_outer[8s + r].  A lane owns 4 consecutive i (one u32 load per
// stripe); per stripe the 8 eq values become two 16-entry nibble tables, built
// LC_CHUNK stripes at a time (one coalesced eq load, one barrier) so that the
// lookups of a chunk run back to back with their z loads in flight.  (16 i per
// lane, one 16-byte load per stripe, measured 2.6x slower at m = 32.)
// A work-group is LC_GROUPS stripe groups of LC_I lanes over the same i: each
// group folds its own quarter of the work-group's stripes, and the groups are
// summed in LDS before the device-scope atomics, which cost about as much as
// the lookups when every 256 lanes issued their own (m = 32: 2048 work-groups
// of 256 lanes 0.52 ms, 4096 0.80 ms, 8192 1.39 ms).
// grid = (k / (4 * LC_I), stripe splits); out must be zeroed.
// ---------------------------------------------------------------------------
constexpr int LC_I = 256;
constexpr int LC_GROUPS = 4;
constexpr int LC_WG = LC_I * LC_GROUPS;
constexpr int LC_CHUNK = 32;

__global__ void __launch_bounds__(LC_WG) lc_partial_fold_kernel(const uint32_t* __restrict__ z, uint64_t k,
                                                               uint64_t n_stripes, uint64_t stripes_per_split,
                                                               const uint4* __restrict__ eq, uint4* __restrict__ out) {
  __shared__ uint4 eqs[LC_GROUPS][LC_CHUNK * 8];
  // [group][stripe][low-nibble table | high-nibble table]; reused for the
  // cross-group sums at the end
  __shared__ uint4 tab[LC_GROUPS][LC_CHUNK][32];
  static_assert(LC_CHUNK * 32 >= LC_I * 4, "the group sums must fit in one group's tables");
  const uint32_t tid = threadIdx.x, li = tid % LC_I, grp = tid / LC_I;
  const uint64_t i4 = ((uint64_t)blockIdx.x * LC_I + li);  // index of the 4-byte group
  const bool active = i4 * 4 < k;
  // This group's stripes: quarter grp of the split.
  const uint64_t sp0 = (uint64_t)blockIdx.y * stripes_per_split;
  uint64_t sp1 = sp0 + stripes_per_split;
  if (sp1 > n_stripes) sp1 = n_stripes;
  const uint64_t per_g = (stripes_per_split + LC_GROUPS - 1) / LC_GROUPS;
  uint64_t s0 = sp0 + grp * per_g, s1 = s0 + per_g;
  if (s0 > sp1) s0 = sp1;
  if (s1 > sp1) s1 = sp1;
  const uint64_t k4 = k / 4;
  uint4 acc[4];
#pragma unroll
  for (int b = 0; b < 4; ++b) acc[b] = f_zero();
  // Every group runs the same number of chunk iterations (barriers are
  // work-group wide); a group whose stripes are used up does empty chunks.
  const uint64_t n_iter = (per_g + LC_CHUNK - 1) / LC_CHUNK;
  uint4* const eg = eqs[grp];
  uint4(*const tg)[32] = tab[grp];
  for (uint64_t it = 0; it < n_iter; ++it) {
    const uint64_t c0 = s0 + it * LC_CHUNK;
    const uint32_t nc = c0 >= s1 ? 0u : (uint32_t)(s1 - c0 < (uint64_t)LC_CHUNK ? s1 - c0 : (uint64_t)LC_CHUNK);
    __syncthreads();  // the previous chunk's lookups are done
    for (uint32_t e = li; e < nc * 8; e += LC_I) eg[e] = eq[8 * c0 + e];
    __syncthreads();
    for (uint32_t e = li; e < nc * 32; e += LC_I) {
      const uint32_t sc = e >> 5, half = (e >> 4) & 1u, kk = e & 15u;
      const uint4* src = eg + sc * 8 + 4 * half;
      uint4 v = f_zero();
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if ((kk >> r) & 1u) v = f_add(v, src[r]);
      tg[sc][e & 31u] = v;
    }
    __syncthreads();
    if (active) {
      const uint32_t* zc = z + c0 * k4 + i4;
#pragma unroll 8
      for (uint32_t sc = 0; sc < nc; ++sc) {
        const uint32_t w = zc[(uint64_t)sc * k4];
        const uint4* t = tg[sc];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const uint32_t byte = (w >> (8 * b)) & 255u;
          const uint4 lo = t[byte & 15u], hi = t[16 + (byte >> 4)];
          acc[b] = make_uint4(xor3(acc[b].x, lo.x, hi.x), xor3(acc[b].y, lo.y, hi.y), xor3(acc[b].z, lo.z, hi.z),
                              xor3(acc[b].w, lo.w, hi.w));
        }
      }
    }
  }
  // Sum the groups: groups 1.. write their sums into their own tables, group 0
  // adds them and issues the atomics.
  __syncthreads();
  if (grp > 0) {
    uint4* red = &tab[grp][0][0];
#pragma unroll
    for (int b = 0; b < 4; ++b) red[b * LC_I + li] = acc[b];
  }
  __syncthreads();
  if (grp == 0 && active) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      uint4 v = acc[b];
#pragma unroll
      for (int g = 1; g < LC_GROUPS; ++g) v = f_add(v, (&tab[g][0][0])[b * LC_I + li]);
      atomic_xor4(out + i4 * 4 + b, v);
    }
  }
}

// ---------------------------------------------------------------------------
// Lincheck product sumcheck, top variable: message and binding.
// ---------------------------------------------------------------------------
constexpr int RD_WG = 128;
// Work-groups of the product-sum kernels (their sums go through partials, so
// the count costs no atomics): FLOCK_RD_GRID overrides it for sweeps.
static unsigned rd_grid_cap() {
  static const unsigned c = [] {
    const char* v = getenv("FLOCK_RD_GRID");
    const long k = v ? atol(v) : 8192;
    return (unsigned)(k < 1 ? 1 : (k > 65535 ? 65535 : k));
  }();
  return c;
}
#define RD_GRID_CAP rd_grid_cap()

__global__ void __launch_bounds__(RD_WG) lc_round_msg_msb_kernel(const uint4* __restrict__ comb,
                                                                const uint4* __restrict__ z, uint64_t half,
                                                                ReduceTail rt) {
  __shared__ uint4 tabs[8 * RD_WG];
  __shared__ uint4 red[2][RD_WG / 64];
  U256 e1, einf;
  u256_zero(e1);
  u256_zero(einf);
  uint4* my = tabs + threadIdx.x;
  table3_init(tabs, threadIdx.x);
  for (uint64_t i = (uint64_t)blockIdx.x * RD_WG + threadIdx.x; i < half; i += (uint64_t)gridDim.x * RD_WG) {
    const uint4 lo = comb[i], hi = comb[i + half], zl = z[i], zh = z[i + half];
    table3_store_private(tabs, RD_WG, threadIdx.x, hi);
    comb3_mul_acc_rt(zh, my, RD_WG, e1);
    table3_store_private(tabs, RD_WG, threadIdx.x, f_add(hi, lo));
    comb3_mul_acc_rt(f_add(zh, zl), my, RD_WG, einf);
  }
  wg_reduce_store(f_reduce(e1), red[0], rt.parts, 0);
  wg_reduce_store(f_reduce(einf), red[1], rt.parts, 1);
}

__global__ void __launch_bounds__(256) lc_bind_top_kernel(uint4* __restrict__ v, uint64_t half, uint4 r) {
  __shared__ uint4 rt[128];  // nibble-position tables of r
  __shared__ uint4 rnib[32];
  build_pos_tables(rt, rnib, 1, [&](uint32_t) { return r; });
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < half; i += (uint64_t)gridDim.x * 256) {
    const uint4 lo = v[i], hi = v[i + half];
    v[i] = combjr_mul_add(f_add(hi, lo), lo, rt);
  }
}

// ---------------------------------------------------------------------------
// Ring switch fold_1b_rows: s_hat_v[r] = sum_i bit_r(z_i) eq_i, r < 128.
// A tile of 256 elements is staged in LDS; every group of 4 elements gets a
// 16-entry subset-sum table of its eq values.  The tile's bits are transposed
// once into nibbles nib[g][r] = (bit r of the group's 4 elements), packed
// eight to a word, so lane (r, half) reads the words of four groups with one
// 16-byte read and looks each group's nibble up.  out: work-group partials, out[r * gridDim.x + blockIdx.x].
// ---------------------------------------------------------------------------
constexpr int F1_WG = 256;
// Entries between the 16-entry subset-sum tables of consecutive element groups:
// the pad of one entry shifts each group's table by one bank group, so the
// quads' table stores (16 groups per wave) spread over the LDS banks instead of
// all hitting the same ones; a lookup still reads one 16-entry table.
constexpr int F1_TS = 17;
// Words per staged element in LDS: 4 plus one of padding, so the transpose's
// reads (lane (g, w) reads word w of the group's four elements) fall in 64
// distinct banks.
constexpr int F1_ZS = 5;
// Words per row of the nibble words, stored [r / 8][group]: 64 groups plus 4 of
// padding, so a wave's 16-byte reads of 4 consecutive groups at its 8 rows fall
// in distinct banks, and so do the transpose's stores (row 4w + j, group g).
constexpr int F1_NP = 68;

// Lane K of the caller's aligned lane quad, broadcast to the quad (DPP).
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t x) {
  constexpr int ctrl = K | (K << 2) | (K << 4) | (K << 6);
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, ctrl, 0xF, 0xF, false);
}
template <int K>
__device__ __forceinline__ uint4 quad_bcast4(uint4 v) {
  return make_uint4(quad_bcast<K>(v.x), quad_bcast<K>(v.y), quad_bcast<K>(v.z), quad_bcast<K>(v.w));
}

// A 4 x 8 bit transpose inside one word: byte k's bit b (bit 8k + b) moves to
// bit 4b + k.  Four delta swaps exchange the index bits (4,2), (3,1), (2,0) and
// (1,0), which turns (k1 k0 b2 b1 b0) into (b2 b1 b0 k1 k0).
__device__ __forceinline__ uint32_t delta_swap(uint32_t x, uint32_t d, uint32_t m) {
  const uint32_t t = __builtin_amdgcn_bitop3_b32(x >> d, x, m, 0x28);  // (x >> d ^ x) & m
  return xor3(x, t, t << d);
}
__device__ __forceinline__ uint32_t transpose_4x8(uint32_t x) {
  x = delta_swap(x, 12, 0x0000F0F0u);
  x = delta_swap(x, 6, 0x00CC00CCu);
  x = delta_swap(x, 3, 0x0A0A0A0Au);
  return delta_swap(x, 1, 0x22222222u);
}

// One tile of 256 elements (z value zv and eq value ev per lane) folded into
// the lane's accumulator acc (output r = t % 128, half t / 128).  Ends with a
// barrier, so the caller may reuse the LDS arrays at once.
__device__ __forceinline__ void f1_tile(uint4 zv, uint4 ev, uint32_t t, uint32_t* zt, uint32_t* nibw, uint4* tabs,
                                        uint4& acc) {
  const uint32_t r = t & 127u, halfsel = t >> 7;
  {
    uint32_t* zr = zt + F1_ZS * t;  // stride 5 words: conflict-free stores and transpose reads
    zr[0] = zv.x;
    zr[1] = zv.y;
    zr[2] = zv.z;
    zr[3] = zv.w;
  }
  // Table of group g = t/4 holds sums over subsets of its 4 eq values; this lane
  // writes the 4 entries whose bits 2,3 equal q = t%4.  The group is the
  // lane's aligned quad, so its eq values arrive by DPP quad broadcasts
  // (VALU moves; cross-lane shuffles would go through the LDS).
  const uint4 e0 = quad_bcast4<0>(ev), e1 = quad_bcast4<1>(ev), e2 = quad_bcast4<2>(ev), e3 = quad_bcast4<3>(ev);
  const uint32_t q = t & 3u;
  uint4 base = f_zero();
  if (q & 1u) base = f_add(base, e2);
  if (q & 2u) base = f_add(base, e3);
  uint4* tg = tabs + (t >> 2) * F1_TS + 4 * q;
  tg[0] = base;
  tg[1] = f_add(base, e0);
  tg[2] = f_add(base, e1);
  tg[3] = f_add(f_add(base, e0), e1);
  __syncthreads();
  // Transpose: word qw = 4w + j of group g packs the nibbles of r = 8qw .. 8qw+7,
  // i.e. byte j of word w of the group's four elements, bit b of element k at
  // bit 4b + k.  Lane t = 4g + w gathers byte j of its four words into one word
  // (byte k = element k) with byte permutes, transposes it in registers, and
  // stores the four words at rows 4w .. 4w+3 of column g.
  {
    const uint32_t g = t >> 2, w = t & 3u;
    const uint32_t* zg = zt + F1_ZS * 4 * g + w;
    const uint32_t e0 = zg[0], e1 = zg[F1_ZS], e2 = zg[2 * F1_ZS], e3 = zg[3 * F1_ZS];
    const uint32_t lo01 = __builtin_amdgcn_perm(e1, e0, 0x05010400u), hi01 = __builtin_amdgcn_perm(e1, e0, 0x07030602u);
    const uint32_t lo23 = __builtin_amdgcn_perm(e3, e2, 0x05010400u), hi23 = __builtin_amdgcn_perm(e3, e2, 0x07030602u);
    uint4 v;
    v.x = transpose_4x8(__builtin_amdgcn_perm(lo23, lo01, 0x05040100u));
    v.y = transpose_4x8(__builtin_amdgcn_perm(lo23, lo01, 0x07060302u));
    v.z = transpose_4x8(__builtin_amdgcn_perm(hi23, hi01, 0x05040100u));
    v.w = transpose_4x8(__builtin_amdgcn_perm(hi23, hi01, 0x07060302u));
    nibw[(4 * w + 0) * F1_NP + g] = v.x;
    nibw[(4 * w + 1) * F1_NP + g] = v.y;
    nibw[(4 * w + 2) * F1_NP + g] = v.z;
    nibw[(4 * w + 3) * F1_NP + g] = v.w;
  }
  __syncthreads();
  const uint32_t qr = r >> 3, shr = 4u * (r & 7u);
  const uint32_t* nrow = nibw + qr * F1_NP;
#pragma unroll 2
  for (uint32_t g = halfsel * 32; g < halfsel * 32 + 32; g += 4) {
    const uint4 nw = *reinterpret_cast<const uint4*>(nrow + g);  // nibble words of groups g .. g+3
    const uint4 x0 = tabs[g * F1_TS + ((nw.x >> shr) & 15u)];
    const uint4 x1 = tabs[(g + 1) * F1_TS + ((nw.y >> shr) & 15u)];
    const uint4 x2 = tabs[(g + 2) * F1_TS + ((nw.z >> shr) & 15u)];
    const uint4 x3 = tabs[(g + 3) * F1_TS + ((nw.w >> shr) & 15u)];
    acc = make_uint4(xor3(acc.x, x0.x, x1.x), xor3(acc.y, x0.y, x1.y), xor3(acc.z, x0.z, x1.z),
                     xor3(acc.w, x0.w, x1.w));
    acc = make_uint4(xor3(acc.x, x2.x, x3.x), xor3(acc.y, x2.y, x3.y), xor3(acc.z, x2.z, x3.z),
                     xor3(acc.w, x2.w, x3.w));
  }
  __syncthreads();
}

__global__ void __launch_bounds__(F1_WG) rs_fold_1b_rows_kernel(const uint4* __restrict__ zp,
                                                               const uint4* __restrict__ eq, uint64_t n,
                                                               uint4* __restrict__ out) {
  __shared__ __align__(16) uint32_t zt[F1_WG * F1_ZS];
  __shared__ __align__(16) uint32_t nibw[16 * F1_NP];  // [r / 8][group]: nibbles of r = 8q .. 8q+7
  __shared__ uint4 tabs[(F1_WG / 4) * F1_TS];
  __shared__ uint4 red[128];
  const uint32_t t = threadIdx.x;
  const uint32_t r = t & 127u, halfsel = t >> 7;
  uint4 acc = f_zero();
  const uint64_t n_tiles = (n + F1_WG - 1) / F1_WG;
  for (uint64_t tile = blockIdx.x; tile < n_tiles; tile += gridDim.x) {
    const uint64_t i = tile * F1_WG + t;
    uint4 zv = f_zero(), ev = f_zero();
    if (i < n) {
      zv = zp[i];
      ev = eq[i];
    }
    f1_tile(zv, ev, t, zt, nibw, tabs, acc);
  }
  if (halfsel == 1) red[r] = acc;
  __syncthreads();
  if (halfsel == 0) out[(size_t)r * gridDim.x + blockIdx.x] = f_add(acc, red[r]);
}

// fold_1b_rows against a split eq tensor eq[i] = eq_lo[i mod 2^lo_log] *
// eq_hi[i >> lo_log] (the factorisation of build_eq_split, CF/pcs/ring_switch.rs:
// 331-357): each block of 2^lo_log elements is folded against eq_lo (read from
// L2, 2^lo_log * 16 bytes), and the block's 128 partial sums are multiplied by
// eq_hi[block] (a 16-entry comb table of it per block) before they are added,
// so the 2^n-element tensor is never written or read.  lo_log >= 8, n a
// multiple of 2^lo_log.
__global__ void __launch_bounds__(F1_WG) rs_fold_1b_rows_split_kernel(const uint4* __restrict__ zp,
                                                                     const uint4* __restrict__ eq_lo,
                                                                     uint32_t lo_log,
                                                                     const uint4* __restrict__ eq_hi,
                                                                     uint64_t n_hi, uint4* __restrict__ out) {
  __shared__ __align__(16) uint32_t zt[F1_WG * F1_ZS];
  __shared__ __align__(16) uint32_t nibw[16 * F1_NP];
  __shared__ uint4 tabs[(F1_WG / 4) * F1_TS];
  __shared__ uint4 red[128];
  __shared__ uint4 etab[16];
  const uint32_t t = threadIdx.x;
  const uint32_t r = t & 127u, halfsel = t >> 7;
  const uint32_t tiles = 1u << (lo_log - 8);
  uint4 tot = f_zero();
  for (uint64_t blk = blockIdx.x; blk < n_hi; blk += gridDim.x) {
    uint4 acc = f_zero();
    const uint4* zb = zp + (blk << lo_log);
    for (uint32_t tt = 0; tt < tiles; ++tt) f1_tile(zb[tt * F1_WG + t], eq_lo[tt * F1_WG + t], t, zt, nibw, tabs, acc);
    if (t < 16) etab[t] = comb_entry(eq_hi[blk], t);
    __syncthreads();
    tot = f_add(tot, comb_mul<1>(acc, etab));
    __syncthreads();
  }
  if (halfsel == 1) red[r] = tot;
  __syncthreads();
  if (halfsel == 0) out[(size_t)r * gridDim.x + blockIdx.x] = f_add(tot, red[r]);
}

// ---------------------------------------------------------------------------
// Ring switch fold_b128_elems: out_i = sum_b bit_b(s_i) eq_r[b].  32 nibble
// tables (table[p][v] = sum of eq_r[4p + bit] over the bits of v, 8 KiB) are
// built once per work-group; each element is 32 lookups.  Byte tables halve
// the lookups, but their data-indexed 4 KiB tables collide in the LDS banks
// (0.38 of the CU cycles at 2^25 elements).
// ---------------------------------------------------------------------------
constexpr int FB_WG = 256;

__global__ void __launch_bounds__(FB_WG) rs_fold_b128_kernel(const uint4* __restrict__ sfx, const uint4* __restrict__ eqr,
                                                            uint64_t n, uint4* __restrict__ out) {
  __shared__ uint4 nibt[32 * 16];
  for (uint32_t e = threadIdx.x; e < 32u * 16u; e += FB_WG) {
    const uint32_t np = e >> 4, k = e & 15u;  // nibble position np covers bits 4np..4np+3
    uint4 v = f_zero();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((k >> j) & 1u) v = f_add(v, eqr[4 * np + j]);
    nibt[e] = v;
  }
  __syncthreads();
  for (uint64_t i = (uint64_t)blockIdx.x * FB_WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * FB_WG) {
    const uint4 s = sfx[i];
    const uint4 r0 = nibtab8_sum(s.x, nibt), r1 = nibtab8_sum(s.y, nibt + 128);
    const uint4 r2 = nibtab8_sum(s.z, nibt + 256), r3 = nibtab8_sum(s.w, nibt + 384);
    out[i] = make_uint4(xor3(r0.x, r1.x, r2.x) ^ r3.x, xor3(r0.y, r1.y, r2.y) ^ r3.y, xor3(r0.z, r1.z, r2.z) ^ r3.z,
                        xor3(r0.w, r1.w, r2.w) ^ r3.w);
  }
}

// ---------------------------------------------------------------------------
// Zerocheck fold at z: out[x] = sum_{s < 2^k_skip} bit(x*2^k_skip + s) * w[s].
// Nibble tables of the weights (2 * row_bytes tables of 16 entries) per
// work-group; 8-byte rows take 16 conflict-free lookups.
// ---------------------------------------------------------------------------
constexpr int FZ_WG = 256;

__global__ void __launch_bounds__(FZ_WG) zc_fold_at_z_kernel(const uint8_t* __restrict__ bits, uint32_t row_bytes,
                                                            uint64_t n_rows, const uint4* __restrict__ w,
                                                            uint4* __restrict__ out) {
  __shared__ uint4 nibt[16 * 16];  // row_bytes (<= 8) * 2 nibble tables
  for (uint32_t e = threadIdx.x; e < row_bytes * 32u; e += FZ_WG) {
    const uint32_t np = e >> 4, k = e & 15u;
    uint4 v = f_zero();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((k >> j) & 1u) v = f_add(v, w[4 * np + j]);
    nibt[e] = v;
  }
  __syncthreads();
  for (uint64_t x = (uint64_t)blockIdx.x * FZ_WG + threadIdx.x; x < n_rows; x += (uint64_t)gridDim.x * FZ_WG) {
    const uint8_t* row = bits + x * row_bytes;
    uint4 acc;
    if (row_bytes == 8) {
      const uint2 v = *reinterpret_cast<const uint2*>(row);
      acc = f_add(nibtab8_sum(v.x, nibt), nibtab8_sum(v.y, nibt + 128));
    } else {
      acc = f_zero();
      for (uint32_t b = 0; b < row_bytes; ++b)
        acc = f_add(acc, f_add(nibt[(2 * b) * 16 + (row[b] & 15u)], nibt[(2 * b + 1) * 16 + (row[b] >> 4)]));
    }
    out[x] = acc;
  }
}

// ---------------------------------------------------------------------------
// Zerocheck multilinear round (wire form): G(1) = sum_x eq[x] a1 b1,
// G(inf) = sum_x eq[x] (a0+a1)(b0+b1), pairs (2x, 2x+1).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RD_WG) zc_round_pair_kernel(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                             const uint4* __restrict__ eq, uint64_t half,
                                                             uint4* __restrict__ out) {
  __shared__ uint4 tabs[8 * RD_WG];
  __shared__ uint4 red[2][RD_WG / 64];
  U256 g1, ginf;
  u256_zero(g1);
  u256_zero(ginf);
  uint4* my = tabs + threadIdx.x;
  table3_init(tabs, threadIdx.x);
  for (uint64_t x = (uint64_t)blockIdx.x * RD_WG + threadIdx.x; x < half; x += (uint64_t)gridDim.x * RD_WG) {
    const uint4 a0 = a[2 * x], a1 = a[2 * x + 1], b0 = b[2 * x], b1 = b[2 * x + 1], e = eq[x];
    table3_store_private(tabs, RD_WG, threadIdx.x, e);
    const uint4 ea1 = comb3_mul_rt(a1, my, RD_WG);
    const uint4 eas = comb3_mul_rt(f_add(a0, a1), my, RD_WG);
    table3_store_private(tabs, RD_WG, threadIdx.x, ea1);
    comb3_mul_acc_rt(b1, my, RD_WG, g1);
    table3_store_private(tabs, RD_WG, threadIdx.x, eas);
    comb3_mul_acc_rt(f_add(b0, b1), my, RD_WG, ginf);
  }
  wg_reduce_store(f_reduce(g1), red[0], out, 0);
  wg_reduce_store(f_reduce(ginf), red[1], out, 1);
}

// ---------------------------------------------------------------------------
// Zerocheck multilinear round with the bind of the previous round fused in and
// the eq table split: T' = fold(T, rho) (when FOLD; T' = T otherwise) is written
// and the wire message over T' is summed, G(1) = sum_x eq[x] a1 b1 and
// G(inf) = sum_x eq[x] (a0+a1)(b0+b1) on the pairs (2x, 2x+1) of T', with
// eq[x] = e_lo[x mod 256] * e_hi[x >> 8].  A work-group iteration takes the 256
// x of one chunk x >> 8, so e_hi[chunk] is one value for the work-group (its
// nibble-position tables are built once per chunk) and e_lo[lane] one value per
// lane, applied once to the lane's sums at the end.  No eq table is written.
// ---------------------------------------------------------------------------
constexpr int ZF_WG = 256;
enum ZfMode { ZF_PLAIN = 0, ZF_FOLD = 1, ZF_BITS = 2 };

// MODE: ZF_PLAIN reads T; ZF_FOLD reads T and writes fold(T, rho); ZF_BITS reads
// the 64-bit rows of the a and b bit vectors and writes T = fold_at_z(bits, w)
// (the weights' nibble tables, w_nib, built per work-group).  eq_hi_next (when
// not null) receives eq_hi with its lowest coordinate summed out, the next
// round's eq_hi: eq_hi_next[y] = eq_hi[2y] + eq_hi[2y + 1].
template <int MODE>
__global__ void __launch_bounds__(ZF_WG) __attribute__((amdgpu_waves_per_eu(4))) zc_fold_round_kernel(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                             uint64_t n_chunks, uint4 rho,
                                                             const uint4* __restrict__ w,
                                                             const uint4* __restrict__ e_lo,
                                                             const uint4* __restrict__ e_hi, uint4* __restrict__ a_out,
                                                             uint4* __restrict__ b_out, uint4* __restrict__ e_hi_next,
                                                             ReduceTail rt) {
  __shared__ uint4 tabs[8 * ZF_WG];  // per-lane tables of the data x data products (3-bit comb)
  __shared__ uint4 rtab[128], rnib[32];
  __shared__ uint4 htab[128], hnib[32];
  __shared__ uint4 wnib[MODE == ZF_BITS ? 16 * 16 : 1];  // 16 nibble positions x 16 subset sums of w
  __shared__ uint4 red[2][ZF_WG / 64];
  const uint32_t tid = threadIdx.x;
  if (MODE == ZF_FOLD) build_pos_tables(rtab, rnib, 1, [&](uint32_t) { return rho; });
  if (MODE == ZF_BITS) {
    const uint32_t np = tid >> 4, k = tid & 15u;  // 256 lanes = 16 positions x 16 values
    uint4 v = f_zero();
#pragma unroll
    for (int j = 0; j < 4; ++j)
      if ((k >> j) & 1u) v = f_add(v, w[4 * np + j]);
    wnib[tid] = v;
    __syncthreads();
  }
  U256 g1, gi;
  u256_zero(g1);
  u256_zero(gi);
  uint4* my = tabs + tid;
  table3_init(tabs, tid);
  // The chunk's inputs are loaded one iteration ahead (u: a, v: b; FOLD 4 each,
  // PLAIN 2 each, BITS 1 each), so their latency overlaps the previous chunk's
  // products and this chunk's table build.
  constexpr int NL = MODE == ZF_FOLD ? 4 : (MODE == ZF_PLAIN ? 2 : 1);
  uint4 u[NL], v[NL];
  auto load = [&](uint64_t c) {
    const uint64_t x = c * ZF_WG + tid;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      u[i] = a[NL * x + i];
      v[i] = b[NL * x + i];
    }
  };
  if (blockIdx.x < n_chunks) load(blockIdx.x);
  for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const uint4 eh = e_hi[c];
    if (e_hi_next && tid == 0 && !(c & 1)) e_hi_next[c >> 1] = f_add(eh, e_hi[c + 1]);
    uint4 cu[NL], cv[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      cu[i] = u[i];
      cv[i] = v[i];
    }
    if (c + gridDim.x < n_chunks) load(c + gridDim.x);
    build_pos_tables(htab, hnib, 1, [&](uint32_t) { return eh; });  // two barriers: the last chunk's lookups are done
    const uint64_t x = c * ZF_WG + tid;
    uint4 a0, a1, b0, b1;
    if (MODE == ZF_FOLD) {
      a0 = combjr_mul_add(f_add(cu[0], cu[1]), cu[0], rtab);  // T'[j] = T[2j] + rho (T[2j] + T[2j+1])
      a1 = combjr_mul_add(f_add(cu[2], cu[3]), cu[2], rtab);
      b0 = combjr_mul_add(f_add(cv[0], cv[1]), cv[0], rtab);
      b1 = combjr_mul_add(f_add(cv[2], cv[3]), cv[2], rtab);
    } else if (MODE == ZF_BITS) {
      // rows 2x and 2x + 1 of each bit vector: 16 bytes = one uint4
      const uint4 ra = cu[0], rb = cv[0];
      a0 = f_add(nibtab8_sum(ra.x, wnib), nibtab8_sum(ra.y, wnib + 128));
      a1 = f_add(nibtab8_sum(ra.z, wnib), nibtab8_sum(ra.w, wnib + 128));
      b0 = f_add(nibtab8_sum(rb.x, wnib), nibtab8_sum(rb.y, wnib + 128));
      b1 = f_add(nibtab8_sum(rb.z, wnib), nibtab8_sum(rb.w, wnib + 128));
    } else {
      a0 = cu[0];
      a1 = cu[1];
      b0 = cv[0];
      b1 = cv[1];
    }
    if (MODE != ZF_PLAIN) {
      a_out[2 * x] = a0;
      a_out[2 * x + 1] = a1;
      b_out[2 * x] = b0;
      b_out[2 * x + 1] = b1;
    }
    const uint4 ea1 = combjr_mul_add(a1, f_zero(), htab);
    const uint4 eas = combjr_mul_add(f_add(a0, a1), f_zero(), htab);
    table3_store_private(tabs, ZF_WG, tid, ea1);
    comb3_mul_acc_rt(b1, my, ZF_WG, g1);
    table3_store_private(tabs, ZF_WG, tid, eas);
    comb3_mul_acc_rt(f_add(b0, b1), my, ZF_WG, gi);
  }
  table3_store_private(tabs, ZF_WG, tid, e_lo[tid]);  // the lane's own table: no barrier needed
  const uint4 r1 = comb3_mul_rt(f_reduce(g1), my, ZF_WG);
  const uint4 ri = comb3_mul_rt(f_reduce(gi), my, ZF_WG);
  wg_reduce_store(r1, red[0], rt.parts, 0);
  wg_reduce_store(ri, red[1], rt.parts, 1);
}

// ---------------------------------------------------------------------------
// The first two multilinear rounds straight from the 64-bit rows of the a and b
// bit vectors, with no table of the fold at z written (the reference's
// uni_skip_fold_and_round_pair over its UniSkipFoldTable,
// CF/zerocheck/multilinear.rs:659-683,744-790).  fold_at_z is GF(2)-linear in a
// row's bits: T[x] = sum_s w_s a(s, x) = N(row x), 16 lookups in the nibble
// tables of the weights (nib[16 p + k] = sum of w_{4p+j} over the bits j of k).
//
// ZB_MSG (round 0): the message over T against eq[x] = eq_lo[x mod 256] *
//   eq_hi[x >> 8].  eq_hi[c] is one value per chunk c of 256 pairs; it goes into
//   the weights of the b side instead of into each pair: chunk c reads its b rows
//   through the nibble tables of eq_hi[c] * w_s (ew[64 c + s],
//   zc_scaled_weights_kernel), so G(1) = sum_x eq_lo * a[2x+1] * (eq_hi b)[2x+1]
//   takes no eq product per pair, and eq_lo is applied per lane at the end.
// ZB_FOLD (round 1): T' = fold(T, rho) from the rows, T'[y] = T[2y] + rho (T[2y] +
//   T[2y+1]) = N(row 2y) + N_rho(row 2y ^ row 2y+1), with N_rho the nibble tables
//   of rho * w_s; T' is written and its message summed as zc_fold_round_kernel
//   sums it (eq_hi through the chunk's nibble-position tables).
// ---------------------------------------------------------------------------
enum ZbMode { ZB_MSG = 0, ZB_FOLD = 1 };

// Entry e (< 256) of the nibble subset-sum tables of the 64 values v[s]:
// XOR of v[4p + j] over the bits j of k, p = e >> 4, k = e & 15.
__device__ __forceinline__ uint4 nib_entry(const uint4* v, uint32_t e) {
  const uint32_t p = e >> 4, k = e & 15u;
  uint4 r = f_zero();
#pragma unroll
  for (int j = 0; j < 4; ++j)
    if ((k >> j) & 1u) r = f_add(r, v[4 * p + j]);
  return r;
}

// The sum selected by one 64-bit row (low word lo: nibbles 0..7, hi: 8..15).
__device__ __forceinline__ uint4 row_sum(uint32_t lo, uint32_t hi, const uint4* nib) {
  return f_add(nibtab8_sum(lo, nib), nibtab8_sum(hi, nib + 128));
}

// ew[64 c + s] = e_hi[c] * w[s] for c < n_hi, s < 64.  Work-group (x, y) builds the
// nibble-position tables of w[16 y .. 16 y + 16); wave v multiplies by weight
// 16 y + 4 k + v (k = 0..3), one e_hi value per lane, so the 64 lanes of a lookup
// share one table (conflict-free).
__global__ void __launch_bounds__(256) zc_scaled_weights_kernel(const uint4* __restrict__ w,
                                                                const uint4* __restrict__ e_hi, uint64_t n_hi,
                                                                uint4* __restrict__ ew) {
  __shared__ uint4 tab[16 * 128], nib[16 * 32];
  const uint32_t s0 = blockIdx.y * 16u, wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
  build_pos_tables(tab, nib, 16, [&](uint32_t c) { return w[s0 + c]; });
  for (uint64_t c = (uint64_t)blockIdx.x * 64 + lane; c < n_hi; c += (uint64_t)gridDim.x * 64) {
    const uint4 e = e_hi[c];
#pragma unroll
    for (uint32_t k = 0; k < 4; ++k) {
      const uint32_t s = 4 * k + wave;
      ew[c * 64 + s0 + s] = combjr_mul_add(e, f_zero(), tab + s * 128);
    }
  }
}

template <int MODE>
__global__ void __launch_bounds__(ZF_WG) __attribute__((amdgpu_waves_per_eu(MODE == ZB_FOLD ? 3 : 2))) zc_bits_round_kernel(const uint4* __restrict__ a, const uint4* __restrict__ b,
                                                             uint64_t n_chunks, uint4 rho,
                                                             const uint4* __restrict__ w,
                                                             const uint4* __restrict__ ew,
                                                             const uint4* __restrict__ e_lo,
                                                             const uint4* __restrict__ e_hi, uint4* __restrict__ a_out,
                                                             uint4* __restrict__ b_out, uint4* __restrict__ e_hi_next,
                                                             ReduceTail rt) {
  __shared__ uint4 tabs[8 * ZF_WG];  // per-lane tables of the data x data products (3-bit comb)
  __shared__ uint4 wnib[256];         // N: nibble tables of w
  __shared__ uint4 xnib[256];         // ZB_MSG: of eq_hi[c] * w (per chunk); ZB_FOLD: of rho * w (N_rho)
  __shared__ uint4 htab[MODE == ZB_FOLD ? 128 : 1], hnib[MODE == ZB_FOLD ? 32 : 1];
  __shared__ uint4 red[2][ZF_WG / 64];
  const uint32_t tid = threadIdx.x;
  wnib[tid] = nib_entry(w, tid);
  if (MODE == ZB_FOLD) {
    // rho * w_s through rho's nibble-position tables (htab is rebuilt per chunk
    // afterwards), staged in the per-lane table area before its first use
    build_pos_tables(htab, hnib, 1, [&](uint32_t) { return rho; });
    if (tid < 64) tabs[tid] = combjr_mul_add(w[tid], f_zero(), htab);
    __syncthreads();
    xnib[tid] = nib_entry(tabs, tid);
  }
  __syncthreads();
  U256 g1, gi;
  u256_zero(g1);
  u256_zero(gi);
  uint4* my = tabs + tid;
  table3_init(tabs, tid);
  // A pair's rows are loaded one chunk ahead: ZB_MSG rows 2x, 2x + 1 (one uint4 per
  // vector), ZB_FOLD rows 4y .. 4y + 3 (two).
  constexpr int NL = MODE == ZB_FOLD ? 2 : 1;
  uint4 u[NL], v[NL];
  auto load = [&](uint64_t c) {
    const uint64_t x = c * ZF_WG + tid;
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      u[i] = a[NL * x + i];
      v[i] = b[NL * x + i];
    }
  };
  if (blockIdx.x < n_chunks) load(blockIdx.x);
  for (uint64_t c = blockIdx.x; c < n_chunks; c += gridDim.x) {
    const uint4 eh = e_hi[c];
    if (e_hi_next && tid == 0 && !(c & 1)) e_hi_next[c >> 1] = f_add(eh, e_hi[c + 1]);
    uint4 cu[NL], cv[NL];
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      cu[i] = u[i];
      cv[i] = v[i];
    }
    if (c + gridDim.x < n_chunks) load(c + gridDim.x);
    const uint64_t x = c * ZF_WG + tid;
    uint4 a1, as, b1, bs;
    if (MODE == ZB_MSG) {
      const uint4 xe = nib_entry(ew + c * 64, tid);  // this lane's entry of the chunk's b tables
      __syncthreads();  // the previous chunk's lookups in xnib are done
      xnib[tid] = xe;
      __syncthreads();
      const uint4 ra = cu[0], rb = cv[0];  // rows 2x (x, y words) and 2x + 1 (z, w)
      a1 = row_sum(ra.z, ra.w, wnib);
      as = row_sum(ra.x ^ ra.z, ra.y ^ ra.w, wnib);
      b1 = row_sum(rb.z, rb.w, xnib);
      bs = row_sum(rb.x ^ rb.z, rb.y ^ rb.w, xnib);
    } else {
      build_pos_tables(htab, hnib, 1, [&](uint32_t) { return eh; });  // two barriers: the last chunk's lookups are done
      const uint4 a0 = f_add(row_sum(cu[0].x, cu[0].y, wnib), row_sum(cu[0].x ^ cu[0].z, cu[0].y ^ cu[0].w, xnib));
      const uint4 ta1 = f_add(row_sum(cu[1].x, cu[1].y, wnib), row_sum(cu[1].x ^ cu[1].z, cu[1].y ^ cu[1].w, xnib));
      const uint4 b0 = f_add(row_sum(cv[0].x, cv[0].y, wnib), row_sum(cv[0].x ^ cv[0].z, cv[0].y ^ cv[0].w, xnib));
      b1 = f_add(row_sum(cv[1].x, cv[1].y, wnib), row_sum(cv[1].x ^ cv[1].z, cv[1].y ^ cv[1].w, xnib));
      a_out[2 * x] = a0;
      a_out[2 * x + 1] = ta1;
      b_out[2 * x] = b0;
      b_out[2 * x + 1] = b1;
      bs = f_add(b0, b1);
      a1 = combjr_mul_add(ta1, f_zero(), htab);
      as = combjr_mul_add(f_add(a0, ta1), f_zero(), htab);
    }
    // four comb steps unrolled: more table loads in flight at two waves per SIMD
    table3_store_private(tabs, ZF_WG, tid, a1);
    comb3_mul_acc_rt<4>(b1, my, ZF_WG, g1);
    table3_store_private(tabs, ZF_WG, tid, as);
    comb3_mul_acc_rt<4>(bs, my, ZF_WG, gi);
  }
  table3_store_private(tabs, ZF_WG, tid, e_lo[tid]);  // the lane's own table: no barrier needed
  const uint4 r1 = comb3_mul_rt(f_reduce(g1), my, ZF_WG);
  const uint4 ri = comb3_mul_rt(f_reduce(gi), my, ZF_WG);
  wg_reduce_store(r1, red[0], rt.parts, 0);
  wg_reduce_store(ri, red[1], rt.parts, 1);
}

static unsigned grid_for(uint64_t items, int wg, unsigned cap) {
  uint64_t g = (items + wg - 1) / wg;
  if (g > cap) g = cap;
  return (unsigned)(g ? g : 1);
}

static inline uint4 u4(flk_f128 a) {
  return make_uint4((uint32_t)a.lo, (uint32_t)(a.lo >> 32), (uint32_t)a.hi, (uint32_t)(a.hi >> 32));
}
static inline uint4 f_zero_host() { return make_uint4(0u, 0u, 0u, 0u); }

int rs_fold_1b_rows_split(const flk_f128* z, const flk_f128* eq_lo, size_t lo_log, const flk_f128* eq_hi, size_t hi_log,
                          flk_f128* out128, hipStream_t st) {
  if (lo_log < 8 || lo_log + hi_log > 40) return set_error(FLK_E_SHAPE, "rs_fold_1b_rows_split: lo_log=%zu", lo_log);
  const uint64_t n_hi = uint64_t(1) << hi_log;
  const unsigned grid = (unsigned)std::min<uint64_t>(n_hi, 2048);
  uint4* o = reinterpret_cast<uint4*>(out128);
  uint4* parts = nullptr;
  int rc = partials_alloc(&parts, grid, 128, o, st);
  if (rc) return rc;
  hipLaunchKernelGGL(rs_fold_1b_rows_split_kernel, dim3(grid), dim3(F1_WG), 0, st, reinterpret_cast<const uint4*>(z),
                     reinterpret_cast<const uint4*>(eq_lo), (uint32_t)lo_log, reinterpret_cast<const uint4*>(eq_hi),
                     n_hi, parts);
  FLK_LAUNCH_CHECK("rs_fold_1b_rows_split_kernel");
  return partials_sum_free(parts, grid, 128, o, st);
}

}  // namespace flk

using namespace flk;

extern "C" {

int flk_lc_partial_fold(const uint8_t* z_packed, size_t m, size_t k_log, const flk_f128* eq_outer, flk_f128* out,
                        void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!z_packed || !eq_outer || !out) return set_error(FLK_E_SHAPE, "lc_partial_fold: null pointer");
  if (k_log < 2 || k_log > m || m - k_log < 3 || m > 40)
    return set_error(FLK_E_SHAPE, "lc_partial_fold: need 2 <= k_log and m - k_log >= 3 (m=%zu k_log=%zu)", m, k_log);
  if ((uintptr_t)z_packed % 4) return set_error(FLK_E_SHAPE, "lc_partial_fold: z_packed must be 4-byte aligned");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const uint64_t k = uint64_t(1) << k_log;
  const uint64_t n_stripes = (uint64_t(1) << (m - k_log)) / 8;
  uint4* o = reinterpret_cast<uint4*>(out);
  if ((rc = zero_n(o, k, st))) return rc;
  const uint64_t gx = (k / 4 + LC_I - 1) / LC_I;
  // ~512 work-groups of 1024 lanes (2 per CU), each split folded by 4 groups
  uint64_t splits = (512 + gx - 1) / gx;
  if (splits > n_stripes) splits = n_stripes;
  if (splits > 65535) splits = 65535;
  const uint64_t per = (n_stripes + splits - 1) / splits;
  splits = (n_stripes + per - 1) / per;
  hipLaunchKernelGGL(lc_partial_fold_kernel, dim3((unsigned)gx, (unsigned)splits), dim3(LC_WG), 0, st,
                     reinterpret_cast<const uint32_t*>(z_packed), k, n_stripes, per,
                     reinterpret_cast<const uint4*>(eq_outer), o);
  FLK_LAUNCH_CHECK("lc_partial_fold_kernel");
  return FLK_OK;
}

int flk_lc_round_msg_msb(const flk_f128* comb, const flk_f128* z, size_t n, flk_f128* out2, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!comb || !z || !out2) return set_error(FLK_E_SHAPE, "lc_round_msg_msb: null pointer");
  if (!is_pow2(n) || n < 2) return set_error(FLK_E_SHAPE, "lc_round_msg_msb: n=%zu must be a power of two >= 2", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out2);
  const unsigned grid = grid_for(n / 2, RD_WG, RD_GRID_CAP);
  ReduceTail rt;
  if ((rc = reduce_begin(grid, 2, o, st, &rt))) return rc;
  hipLaunchKernelGGL(lc_round_msg_msb_kernel, dim3(grid), dim3(RD_WG), 0, st, reinterpret_cast<const uint4*>(comb),
                     reinterpret_cast<const uint4*>(z), (uint64_t)(n / 2), rt);
  FLK_LAUNCH_CHECK("lc_round_msg_msb_kernel");
  return reduce_end(rt, grid, 2, st);
}

int flk_lc_bind_top(flk_f128* v, size_t n, flk_f128 r, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!v) return set_error(FLK_E_SHAPE, "lc_bind_top: null pointer");
  if (!is_pow2(n) || n < 2) return set_error(FLK_E_SHAPE, "lc_bind_top: n=%zu must be a power of two >= 2", n);
  hipLaunchKernelGGL(lc_bind_top_kernel, dim3(grid_for(n / 2, 256, 65536)), dim3(256), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<uint4*>(v), (uint64_t)(n / 2), u4(r));
  FLK_LAUNCH_CHECK("lc_bind_top_kernel");
  return FLK_OK;
}

int flk_rs_fold_1b_rows(const flk_f128* packed_witness, const flk_f128* suffix_tensor, size_t n, flk_f128* out128,
                        void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!packed_witness || !suffix_tensor || !out128) return set_error(FLK_E_SHAPE, "rs_fold_1b_rows: null pointer");
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out128);
  if (n == 0) return zero_n(o, 128, st);
  const unsigned grid = grid_for(n, F1_WG, 2048);
  uint4* parts = nullptr;
  if ((rc = partials_alloc(&parts, grid, 128, o, st))) return rc;
  hipLaunchKernelGGL(rs_fold_1b_rows_kernel, dim3(grid), dim3(F1_WG), 0, st,
                     reinterpret_cast<const uint4*>(packed_witness), reinterpret_cast<const uint4*>(suffix_tensor),
                     (uint64_t)n, parts);
  FLK_LAUNCH_CHECK("rs_fold_1b_rows_kernel");
  return partials_sum_free(parts, grid, 128, o, st);
}

int flk_rs_fold_b128_elems(const flk_f128* suffix_tensor, const flk_f128* eq_r_dprime, size_t n, flk_f128* out,
                           void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!suffix_tensor || !eq_r_dprime || !out) return set_error(FLK_E_SHAPE, "rs_fold_b128_elems: null pointer");
  if (n == 0) return FLK_OK;
  hipLaunchKernelGGL(rs_fold_b128_kernel, dim3(grid_for(n, FB_WG, 1024)), dim3(FB_WG), 0,
                     static_cast<hipStream_t>(stream), reinterpret_cast<const uint4*>(suffix_tensor),
                     reinterpret_cast<const uint4*>(eq_r_dprime), (uint64_t)n, reinterpret_cast<uint4*>(out));
  FLK_LAUNCH_CHECK("rs_fold_b128_kernel");
  return FLK_OK;
}

int flk_zc_fold_at_z(const uint8_t* bits, size_t m, size_t k_skip, const flk_f128* weights, flk_f128* out,
                     void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!bits || !weights || !out) return set_error(FLK_E_SHAPE, "zc_fold_at_z: null pointer");
  if (k_skip < 3 || k_skip > 6 || k_skip > m || m > 40)
    return set_error(FLK_E_SHAPE, "zc_fold_at_z: k_skip=%zu unsupported (3..6, <= m=%zu)", k_skip, m);
  if (k_skip == 6 && (uintptr_t)bits % 8) return set_error(FLK_E_SHAPE, "zc_fold_at_z: bits must be 8-byte aligned");
  const uint32_t row_bytes = (uint32_t)((size_t(1) << k_skip) / 8);
  const uint64_t n_rows = uint64_t(1) << (m - k_skip);
  hipLaunchKernelGGL(zc_fold_at_z_kernel, dim3(grid_for(n_rows, FZ_WG, 2048)), dim3(FZ_WG), 0,
                     static_cast<hipStream_t>(stream), bits, row_bytes, n_rows, reinterpret_cast<const uint4*>(weights),
                     reinterpret_cast<uint4*>(out));
  FLK_LAUNCH_CHECK("zc_fold_at_z_kernel");
  return FLK_OK;
}

static int zc_fold_round_launch(int mode, const void* a, const void* b, uint64_t hi_log, const flk_f128* rho,
                                const flk_f128* w, const flk_f128* eq_lo, const flk_f128* eq_hi, flk_f128* a_out,
                                flk_f128* b_out, flk_f128* eq_hi_next, flk_f128* out2, hipStream_t st) {
  uint4* o = reinterpret_cast<uint4*>(out2);
  const uint64_t n_chunks = uint64_t(1) << hi_log;
  const unsigned grid = (unsigned)std::min<uint64_t>(n_chunks, 2048);
  ReduceTail rt;
  int rc = reduce_begin(grid, 2, o, st, &rt);
  if (rc) return rc;
  const uint4 r = rho ? u4(*rho) : f_zero_host();
  auto* ka = reinterpret_cast<const uint4*>(a);
  auto* kb = reinterpret_cast<const uint4*>(b);
  auto* kw = reinterpret_cast<const uint4*>(w);
  auto* kl = reinterpret_cast<const uint4*>(eq_lo);
  auto* kh = reinterpret_cast<const uint4*>(eq_hi);
  auto* ao = reinterpret_cast<uint4*>(a_out);
  auto* bo = reinterpret_cast<uint4*>(b_out);
  auto* hn = hi_log ? reinterpret_cast<uint4*>(eq_hi_next) : nullptr;
  if (mode == ZF_FOLD)
    hipLaunchKernelGGL(zc_fold_round_kernel<ZF_FOLD>, dim3(grid), dim3(ZF_WG), 0, st, ka, kb, n_chunks, r, kw, kl, kh,
                       ao, bo, hn, rt);
  else if (mode == ZF_BITS)
    hipLaunchKernelGGL(zc_fold_round_kernel<ZF_BITS>, dim3(grid), dim3(ZF_WG), 0, st, ka, kb, n_chunks, r, kw, kl, kh,
                       ao, bo, hn, rt);
  else
    hipLaunchKernelGGL(zc_fold_round_kernel<ZF_PLAIN>, dim3(grid), dim3(ZF_WG), 0, st, ka, kb, n_chunks, r, kw, kl,
                       kh, ao, bo, hn, rt);
  FLK_LAUNCH_CHECK("zc_fold_round_kernel");
  return reduce_end(rt, grid, 2, st);
}

// The two bits rounds (zc_bits_round_kernel): ZB_MSG takes the eq_hi-scaled
// weights from a scratch table written first by zc_scaled_weights_kernel.
static int zc_bits_round_launch(int mode, const uint8_t* a_bits, const uint8_t* b_bits, uint64_t hi_log,
                                const flk_f128* rho, const flk_f128* w, const flk_f128* eq_lo, const flk_f128* eq_hi,
                                flk_f128* a_out, flk_f128* b_out, flk_f128* eq_hi_next, flk_f128* out2,
                                hipStream_t st) {
  uint4* o = reinterpret_cast<uint4*>(out2);
  const uint64_t n_chunks = uint64_t(1) << hi_log;
  const unsigned grid = (unsigned)std::min<uint64_t>(n_chunks, 2048);
  auto* kw = reinterpret_cast<const uint4*>(w);
  auto* kh = reinterpret_cast<const uint4*>(eq_hi);
  uint4* ew = nullptr;
  if (mode == ZB_MSG) {
    hipError_t e = stream_alloc(reinterpret_cast<void**>(&ew), n_chunks * 64 * sizeof(uint4), st);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return set_error(FLK_E_NOMEM, "zc_fold_at_z_round: scaled-weight scratch of %llu bytes",
                       (unsigned long long)(n_chunks * 64 * sizeof(uint4)));
    }
    const unsigned gx = (unsigned)std::min<uint64_t>((n_chunks + 63) / 64, 1024);
    hipLaunchKernelGGL(zc_scaled_weights_kernel, dim3(gx, 4), dim3(256), 0, st, kw, kh, n_chunks, ew);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) {
      (void)stream_free(ew, st);
      return device_failure(le, "zc_scaled_weights_kernel");
    }
  }
  ReduceTail rt;
  int rc = reduce_begin(grid, 2, o, st, &rt);
  if (rc) {
    if (ew) (void)stream_free(ew, st);
    return rc;
  }
  const uint4 r = rho ? u4(*rho) : f_zero_host();
  auto* ka = reinterpret_cast<const uint4*>(a_bits);
  auto* kb = reinterpret_cast<const uint4*>(b_bits);
  auto* kl = reinterpret_cast<const uint4*>(eq_lo);
  auto* ao = reinterpret_cast<uint4*>(a_out);
  auto* bo = reinterpret_cast<uint4*>(b_out);
  auto* hn = hi_log ? reinterpret_cast<uint4*>(eq_hi_next) : nullptr;
  if (mode == ZB_MSG)
    hipLaunchKernelGGL(zc_bits_round_kernel<ZB_MSG>, dim3(grid), dim3(ZF_WG), 0, st, ka, kb, n_chunks, r, kw, ew, kl,
                       kh, ao, bo, hn, rt);
  else
    hipLaunchKernelGGL(zc_bits_round_kernel<ZB_FOLD>, dim3(grid), dim3(ZF_WG), 0, st, ka, kb, n_chunks, r, kw, ew, kl,
                       kh, ao, bo, hn, rt);
  hipError_t le = hipGetLastError();
  if (le != hipSuccess) {
    if (ew) (void)stream_free(ew, st);
    return device_failure(le, "zc_bits_round_kernel");
  }
  rc = reduce_end(rt, grid, 2, st);
  if (ew) {
    hipError_t fe = stream_free(ew, st);
    if (!rc && fe != hipSuccess) rc = device_failure(fe, "zc_fold_at_z_round: scratch free");
  }
  return rc;
}

int flk_zc_fold_round(const flk_f128* a, const flk_f128* b, size_t n, const flk_f128* rho, const flk_f128* eq_lo,
                      const flk_f128* eq_hi, size_t hi_log, flk_f128* a_out, flk_f128* b_out, flk_f128* eq_hi_next,
                      flk_f128* out2, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!a || !b || !eq_lo || !eq_hi || !out2 || (rho && (!a_out || !b_out)))
    return set_error(FLK_E_SHAPE, "zc_fold_round: null pointer");
  if (hi_log > 32) return set_error(FLK_E_SHAPE, "zc_fold_round: hi_log=%zu", hi_log);
  const uint64_t n_pairs = uint64_t(ZF_WG) << hi_log;  // pairs of the (folded) table
  if (n != (rho ? 4 * n_pairs : 2 * n_pairs))
    return set_error(FLK_E_SHAPE, "zc_fold_round: n=%zu, expected %llu for hi_log=%zu%s", n,
                     (unsigned long long)(rho ? 4 * n_pairs : 2 * n_pairs), hi_log, rho ? " with a fold" : "");
  return zc_fold_round_launch(rho ? ZF_FOLD : ZF_PLAIN, a, b, hi_log, rho, nullptr, eq_lo, eq_hi, a_out, b_out,
                              eq_hi_next, out2, static_cast<hipStream_t>(stream));
}

int flk_zc_fold_at_z_round(const uint8_t* a_bits, const uint8_t* b_bits, size_t m, const flk_f128* weights,
                           const flk_f128* eq_lo, const flk_f128* eq_hi, flk_f128* a_out, flk_f128* b_out,
                           flk_f128* eq_hi_next, flk_f128* out2, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!a_bits || !b_bits || !weights || !eq_lo || !eq_hi || !out2 || (!a_out != !b_out))
    return set_error(FLK_E_SHAPE, "zc_fold_at_z_round: null pointer");
  if (m < 6 + 1 + 8 || m > 40) return set_error(FLK_E_SHAPE, "zc_fold_at_z_round: m=%zu outside 15..40", m);
  if (((uintptr_t)a_bits | (uintptr_t)b_bits) % 16)
    return set_error(FLK_E_SHAPE, "zc_fold_at_z_round: bit vectors must be 16-byte aligned");
  // 2^(m-6) rows of 64 bits; the pairs of rows are 2^(m-7) = 256 * 2^hi_log
  if (!a_out)  // the message alone: nothing written but out2 (and eq_hi_next)
    return zc_bits_round_launch(ZB_MSG, a_bits, b_bits, m - 7 - 8, nullptr, weights, eq_lo, eq_hi, nullptr, nullptr,
                                eq_hi_next, out2, static_cast<hipStream_t>(stream));
  return zc_fold_round_launch(ZF_BITS, a_bits, b_bits, m - 7 - 8, nullptr, weights, eq_lo, eq_hi, a_out, b_out,
                              eq_hi_next, out2, static_cast<hipStream_t>(stream));
}

int flk_zc_fold_at_z_bind_round(const uint8_t* a_bits, const uint8_t* b_bits, size_t m, const flk_f128* weights,
                                const flk_f128* rho, const flk_f128* eq_lo, const flk_f128* eq_hi, flk_f128* a_out,
                                flk_f128* b_out, flk_f128* eq_hi_next, flk_f128* out2, void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!a_bits || !b_bits || !weights || !rho || !eq_lo || !eq_hi || !a_out || !b_out || !out2)
    return set_error(FLK_E_SHAPE, "zc_fold_at_z_bind_round: null pointer");
  if (m < 6 + 2 + 8 || m > 40) return set_error(FLK_E_SHAPE, "zc_fold_at_z_bind_round: m=%zu outside 16..40", m);
  if (((uintptr_t)a_bits | (uintptr_t)b_bits) % 16)
    return set_error(FLK_E_SHAPE, "zc_fold_at_z_bind_round: bit vectors must be 16-byte aligned");
  // T' has 2^(m-7) entries: 2^(m-8) = 256 * 2^hi_log pairs, each from four rows
  return zc_bits_round_launch(ZB_FOLD, a_bits, b_bits, m - 8 - 8, rho, weights, eq_lo, eq_hi, a_out, b_out,
                              eq_hi_next, out2, static_cast<hipStream_t>(stream));
}

int flk_zc_round_pair(const flk_f128* a, const flk_f128* b, size_t n, const flk_f128* eq_rest, flk_f128* out2,
                      void* stream) {
  int rc = check_enabled();
  if (rc) return rc;
  if (!a || !b || !eq_rest || !out2) return set_error(FLK_E_SHAPE, "zc_round_pair: null pointer");
  if (!is_pow2(n) || n < 2) return set_error(FLK_E_SHAPE, "zc_round_pair: n=%zu must be a power of two >= 2", n);
  hipStream_t st = static_cast<hipStream_t>(stream);
  uint4* o = reinterpret_cast<uint4*>(out2);
  const unsigned grid = grid_for(n / 2, RD_WG, RD_GRID_CAP);
  uint4* parts = nullptr;
  if ((rc = partials_alloc(&parts, grid, 2, o, st))) return rc;
  hipLaunchKernelGGL(zc_round_pair_kernel, dim3(grid), dim3(RD_WG), 0, st, reinterpret_cast<const uint4*>(a),
                     reinterpret_cast<const uint4*>(b), reinterpret_cast<const uint4*>(eq_rest), (uint64_t)(n / 2),
                     parts);
  FLK_LAUNCH_CHECK("zc_round_pair_kernel");
  return partials_sum_free(parts, grid, 2, o, st);
}

}  // extern "C"
