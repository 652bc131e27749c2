// BLAKE3 compression for gfx950, written from the BLAKE3 specification
// (the reference uses crate blake3 1.8.6 through its hazmat API,
// CF/merkle.rs:192-203).  One lane runs one compression chain; the seven
// rounds and the message schedule are fully unrolled so every message word is
// a register and each rotation is one v_alignbit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace flk {

enum : uint32_t { B3_CHUNK_START = 1u, B3_CHUNK_END = 2u, B3_PARENT = 4u, B3_ROOT = 8u };

__device__ __constant__ static const uint32_t B3_IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                                                         0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};

__device__ __forceinline__ uint32_t rotr32(uint32_t x, uint32_t n) { return __builtin_amdgcn_alignbit(x, x, n); }

#define FLK_B3_G(a, b, c, d, x, y)   \
  do {                               \
    a = a + b + (x);                 \
    d = rotr32(d ^ a, 16);           \
    c = c + d;                       \
    b = rotr32(b ^ c, 12);           \
    a = a + b + (y);                 \
    d = rotr32(d ^ a, 8);            \
    c = c + d;                       \
    b = rotr32(b ^ c, 7);            \
  } while (0)

// Message word order for round r is m[SCHED_r[i]] with
// SCHED_0 = identity and SCHED_{r+1}[i] = SCHED_r[PERM[i]],
// PERM = {2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8}.
#define FLK_B3_ROUND(m0, m1, m2, m3, m4, m5, m6, m7, m8, m9, m10, m11, m12, m13, m14, m15) \
  do {                                                                                      \
    FLK_B3_G(v0, v4, v8, v12, m[m0], m[m1]);                                                \
    FLK_B3_G(v1, v5, v9, v13, m[m2], m[m3]);                                                \
    FLK_B3_G(v2, v6, v10, v14, m[m4], m[m5]);                                               \
    FLK_B3_G(v3, v7, v11, v15, m[m6], m[m7]);                                               \
    FLK_B3_G(v0, v5, v10, v15, m[m8], m[m9]);                                               \
    FLK_B3_G(v1, v6, v11, v12, m[m10], m[m11]);                                             \
    FLK_B3_G(v2, v7, v8, v13, m[m12], m[m13]);                                              \
    FLK_B3_G(v3, v4, v9, v14, m[m14], m[m15]);                                              \
  } while (0)

// cv <- first 8 words of compress(cv, m, counter, block_len, flags).
__device__ __forceinline__ void b3_compress_cv(uint32_t cv[8], const uint32_t m[16], uint32_t counter_lo,
                                               uint32_t counter_hi, uint32_t block_len, uint32_t flags) {
  uint32_t v0 = cv[0], v1 = cv[1], v2 = cv[2], v3 = cv[3];
  uint32_t v4 = cv[4], v5 = cv[5], v6 = cv[6], v7 = cv[7];
  uint32_t v8 = 0x6A09E667u, v9 = 0xBB67AE85u, v10 = 0x3C6EF372u, v11 = 0xA54FF53Au;
  uint32_t v12 = counter_lo, v13 = counter_hi, v14 = block_len, v15 = flags;
  FLK_B3_ROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  FLK_B3_ROUND(2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8);
  FLK_B3_ROUND(3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1);
  FLK_B3_ROUND(10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6);
  FLK_B3_ROUND(12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4);
  FLK_B3_ROUND(9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7);
  FLK_B3_ROUND(11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13);
  cv[0] = v0 ^ v8;
  cv[1] = v1 ^ v9;
  cv[2] = v2 ^ v10;
  cv[3] = v3 ^ v11;
  cv[4] = v4 ^ v12;
  cv[5] = v5 ^ v13;
  cv[6] = v6 ^ v14;
  cv[7] = v7 ^ v15;
}

__device__ __forceinline__ void b3_iv(uint32_t cv[8]) {
  cv[0] = 0x6A09E667u;
  cv[1] = 0xBB67AE85u;
  cv[2] = 0x3C6EF372u;
  cv[3] = 0xA54FF53Au;
  cv[4] = 0x510E527Fu;
  cv[5] = 0x9B05688Cu;
  cv[6] = 0x1F83D9ABu;
  cv[7] = 0x5BE0CD19u;
}

// Non-root parent chaining value of two children (CF/merkle.rs:198-203).
__device__ __forceinline__ void b3_parent(const uint32_t l[8], const uint32_t r[8], uint32_t out[8]) {
  uint32_t m[16];
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    m[i] = l[i];
    m[8 + i] = r[i];
  }
  b3_iv(out);
  b3_compress_cv(out, m, 0u, 0u, 64u, B3_PARENT);
}

// ---------------------------------------------------------------------------
// Quad-cooperative parent compression for latency-bound tree tops: the four
// lanes of an aligned lane quad (q = lane & 3) each run one column of the
// 4 x 4 state (a = v[q], b = v[4+q], c = v[8+q], d = v[12+q]); for the
// diagonal step rows b, c, d are rotated by 1, 2, 3 lanes with DPP quad
// permutes and rotated back after it.  Every lane must hold the same 16 message
// words (left CV, then right CV) and all four lanes of the quad must be active.
// Lane q returns output words q and 4 + q.
// ---------------------------------------------------------------------------
template <int K>  // lane i reads lane (i + K) mod 4 of its quad
__device__ __forceinline__ uint32_t quad_rot(uint32_t x) {
  constexpr int ctrl = ((0 + K) & 3) | (((1 + K) & 3) << 2) | (((2 + K) & 3) << 4) | (((3 + K) & 3) << 6);
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)x, ctrl, 0xF, 0xF, false);
}

__device__ __forceinline__ uint32_t sel_quad(bool q1, bool q2, uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3) {
  return q2 ? (q1 ? w3 : w2) : (q1 ? w1 : w0);
}

#define FLK_B3_QROUND(s0, s1, s2, s3, s4, s5, s6, s7, s8, s9, s10, s11, s12, s13, s14, s15)            \
  do {                                                                                                 \
    FLK_B3_G(a, b, c, d, sel_quad(q1, q2, m[s0], m[s2], m[s4], m[s6]),                                 \
             sel_quad(q1, q2, m[s1], m[s3], m[s5], m[s7]));                                            \
    b = quad_rot<1>(b);                                                                                \
    c = quad_rot<2>(c);                                                                                \
    d = quad_rot<3>(d);                                                                                \
    FLK_B3_G(a, b, c, d, sel_quad(q1, q2, m[s8], m[s10], m[s12], m[s14]),                              \
             sel_quad(q1, q2, m[s9], m[s11], m[s13], m[s15]));                                         \
    b = quad_rot<3>(b);                                                                                \
    c = quad_rot<2>(c);                                                                                \
    d = quad_rot<1>(d);                                                                                \
  } while (0)

__device__ __forceinline__ uint2 b3_parent_quad(const uint32_t m[16], uint32_t q) {
  const bool q1 = (q & 1u) != 0u, q2 = (q & 2u) != 0u;
  uint32_t a = sel_quad(q1, q2, 0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au);
  uint32_t b = sel_quad(q1, q2, 0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u);
  uint32_t c = a;                                       // v[8..11] = IV[0..3]
  uint32_t d = sel_quad(q1, q2, 0u, 0u, 64u, B3_PARENT);  // counter 0, block_len 64, PARENT
  FLK_B3_QROUND(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
  FLK_B3_QROUND(2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8);
  FLK_B3_QROUND(3, 4, 10, 12, 13, 2, 7, 14, 6, 5, 9, 0, 11, 15, 8, 1);
  FLK_B3_QROUND(10, 7, 12, 9, 14, 3, 13, 15, 4, 0, 11, 2, 5, 8, 1, 6);
  FLK_B3_QROUND(12, 13, 9, 11, 15, 10, 14, 8, 7, 2, 5, 3, 0, 1, 6, 4);
  FLK_B3_QROUND(9, 14, 11, 5, 8, 12, 15, 1, 13, 3, 0, 10, 2, 6, 4, 7);
  FLK_B3_QROUND(11, 15, 5, 0, 1, 9, 8, 6, 14, 10, 2, 12, 3, 4, 7, 13);
  return make_uint2(a ^ c, b ^ d);  // out[q] = v[q] ^ v[8+q], out[4+q] = v[4+q] ^ v[12+q]
}

}  // namespace flk
