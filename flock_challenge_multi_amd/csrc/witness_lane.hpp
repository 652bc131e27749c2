// Per-compression part of the BLAKE3 R1CS witness generator: the row streams
// of one block and the 8-block bit transpose of the lincheck stripe.  Shared by
// the device kernel (witness.hip, one lane per compression) and a host-side
// unit test; nothing here depends on the device.
//
// Reference: CP/r1cs_hashes/blake3.rs (CP = crates/flock-prover/src/).
//   layout                        :111-241   K = 2^14 bits per compression
//   row witnesses                 :1886-1893 (z, a, b) per row kind, C = I
//   build_block_witness_ab_packed_into :2044-2182 (bit order of the packed rows)
//   add_carry_parts               CP/r1cs_hashes/common.rs:80-88
//   lincheck stripe               CP/r1cs_hashes/common.rs:304-320,
//                                 CF/bits.rs:22-67 (CF = crates/flock-core/src/)
//
// Bit p of a block's K-bit row vector is bit p % 64 of u64 word p / 64.  The
// words of a block come out strictly in order 0..255 through Sink::put(widx,
// z, a, b); the caller's Sink::checkpoint(widx) runs after every section and
// after every G call, when at most 19 words are pending since the last
// 16-word boundary that the sink has passed.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#define FLK_WHD __host__ __device__ __forceinline__
#else
#define FLK_WHD inline
#endif

namespace flk {
namespace wit {

constexpr int K_BITS = 1 << 14;
constexpr int WORDS = K_BITS / 64;  // 256 u64 per block and row vector
constexpr int CONST_POS = 512;
constexpr int GS_BASE = 1153;
constexpr int G_STRIDE = 250;
constexpr int OUT_HI_BASE = GS_BASE + 56 * G_STRIDE;  // 15153
constexpr int USEFUL_BITS = OUT_HI_BASE + 256;        // 15409

struct Sched {
  uint8_t v[7][16];
};
// Message word consumed at position j of round r: PERM^r applied to the
// identity (blake3.rs:149,325-340).
constexpr Sched make_sched() {
  const uint8_t perm[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
  Sched s{};
  uint8_t p[16] = {};
  for (int i = 0; i < 16; ++i) p[i] = (uint8_t)i;
  for (int r = 0; r < 7; ++r) {
    for (int i = 0; i < 16; ++i) s.v[r][i] = p[i];
    uint8_t n[16] = {};
    for (int i = 0; i < 16; ++i) n[i] = p[perm[i]];
    for (int i = 0; i < 16; ++i) p[i] = n[i];
  }
  return s;
}

FLK_WHD uint32_t iv(int i) {
  return i == 0 ? 0x6a09e667u : i == 1 ? 0xbb67ae85u : i == 2 ? 0x3c6ef372u : i == 3 ? 0xa54ff53au
       : i == 4 ? 0x510e527fu : i == 5 ? 0x9b05688cu : i == 6 ? 0x1f83d9abu : 0x5be0cd19u;
}

FLK_WHD uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// The compression's state after the 7 rounds (before the finalisation XORs),
// in the textbook lane order; `msg(i)` returns message word i.
template <class Msg>
FLK_WHD void compress_state(const uint32_t cv[8], const Msg& msg, uint32_t t_lo, uint32_t t_hi,
                            uint32_t blen, uint32_t flags, const Sched& sc, uint32_t s[16]) {
  for (int i = 0; i < 8; ++i) s[i] = cv[i];
  for (int i = 0; i < 4; ++i) s[8 + i] = iv(i);
  s[12] = t_lo;
  s[13] = t_hi;
  s[14] = blen;
  s[15] = flags;
  const int lanes[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                           {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};
#pragma unroll
  for (int r = 0; r < 7; ++r) {
#pragma unroll
    for (int g = 0; g < 8; ++g) {
      const int a = lanes[g][0], b = lanes[g][1], c = lanes[g][2], d = lanes[g][3];
      s[a] = s[a] + s[b] + msg(sc.v[r][2 * g]);
      s[d] = rotr32(s[d] ^ s[a], 16);
      s[c] = s[c] + s[d];
      s[b] = rotr32(s[b] ^ s[c], 12);
      s[a] = s[a] + s[b] + msg(sc.v[r][2 * g + 1]);
      s[d] = rotr32(s[d] ^ s[a], 8);
      s[c] = s[c] + s[d];
      s[b] = rotr32(s[b] ^ s[c], 7);
    }
  }
}

// The three row vectors (z, a, b) of one block as a stream of u64 words.
template <class Sink>
struct RowStream {
  Sink& sink;
  uint64_t z, a, b;  // pending low bits of the current word
  int used;          // number of pending bits, < 64
  int widx;          // index of the next word

  FLK_WHD void word(uint64_t zw, uint64_t aw, uint64_t bw) {
    sink.put(widx, zw, aw, bw);
    ++widx;
  }
  // Append W bits of each vector (values already masked to W bits).
  template <int W>
  FLK_WHD void push(uint32_t zv, uint32_t av, uint32_t bv) {
    z |= (uint64_t)zv << used;
    a |= (uint64_t)av << used;
    b |= (uint64_t)bv << used;
    used += W;
    if (used >= 64) {
      word(z, a, b);
      used -= 64;
      const int sh = W - used;  // bits of this value already in the emitted word, 1..W
      z = (uint64_t)zv >> sh;
      a = (uint64_t)av >> sh;
      b = (uint64_t)bv >> sh;
    }
  }
  // A lin row word (blake3.rs:1889-1890): z = a = v, b = 1.
  FLK_WHD void lin(uint32_t v) { push<32>(v, v, 0xffffffffu); }
  // One 32-bit addition: its 31 carry rows (common.rs:80-88) and the sum.
  FLK_WHD uint32_t add(uint32_t x, uint32_t y) {
    const uint32_t sum = x + y;
    const uint32_t cin = sum ^ x ^ y;
    const uint32_t left = (x ^ cin) & 0x7fffffffu;
    const uint32_t right = (y ^ cin) & 0x7fffffffu;
    push<31>(left & right, left, right);
    return sum;
  }
};

// Rotate row `r` of the 4x4 state left by `n` columns (new column j = old
// column j + n).
FLK_WHD void rot_row(uint32_t s[16], int r, int n) {
  uint32_t t[4];
  for (int j = 0; j < 4; ++j) t[j] = s[4 * r + ((j + n) & 3)];
  for (int j = 0; j < 4; ++j) s[4 * r + j] = t[j];
}

// All 256 words of one block's (z, a, b).  out_lo[w] = final state[w] ^
// state[w + 8], from compress_state (its row words precede the G records).
template <class Sink, class Msg>
FLK_WHD void block_rows(const uint32_t cv[8], const Msg& msg, uint32_t t_lo, uint32_t t_hi, uint32_t blen,
                        uint32_t flags, const uint32_t out_lo[8], const Sched& sc, Sink& sink) {
  RowStream<Sink> st{sink, 0, 0, 0, 0, 0};
  const uint64_t ones = ~0ull;
  // words 0..3: cv; 4..7: out_lo (aligned lin rows)
  for (int i = 0; i < 4; ++i) {
    const uint64_t v = (uint64_t)cv[2 * i] | ((uint64_t)cv[2 * i + 1] << 32);
    st.word(v, v, ones);
  }
  for (int i = 0; i < 4; ++i) {
    const uint64_t v = (uint64_t)out_lo[2 * i] | ((uint64_t)out_lo[2 * i + 1] << 32);
    st.word(v, v, ones);
  }
  // words 8..17: the constant wire at bit 512, then m[0..16), t_lo, t_hi,
  // block_len, flags from bit 513 on (input rows: z = a = value, b = 1)
  uint32_t prev_top = 1;  // bit 0 of word 8 is the constant wire
  for (int i = 0; i < 10; ++i) {
    const uint32_t lo = 2 * i < 16 ? msg(2 * i) : (2 * i == 16 ? t_lo : blen);
    const uint32_t hi = 2 * i + 1 < 16 ? msg(2 * i + 1) : (2 * i + 1 == 17 ? t_hi : flags);
    const uint64_t v = (uint64_t)prev_top | ((uint64_t)lo << 1) | ((uint64_t)hi << 33);
    st.word(v, v, ones);
    prev_top = hi >> 31;
  }
  st.z = st.a = prev_top;  // flags bit 31 at bit 1152
  st.b = 1;
  st.used = 1;
  sink.checkpoint(st.widx);

  // The 56 G calls.  The state is kept rotated so that every call works on
  // column 0: rows are rotated left by (g, g, g, g) for column call g and by
  // (j, 1 + j, 2 + j, 3 + j) for diagonal call 4 + j.
  uint32_t s[16];
  for (int i = 0; i < 8; ++i) s[i] = cv[i];
  for (int i = 0; i < 4; ++i) s[8 + i] = iv(i);
  s[12] = t_lo;
  s[13] = t_hi;
  s[14] = blen;
  s[15] = flags;
#pragma unroll 1
  for (int r = 0; r < 7; ++r) {
#pragma unroll 1
    for (int g = 0; g < 8; ++g) {
      const uint32_t mx = msg(sc.v[r][2 * g]), my = msg(sc.v[r][2 * g + 1]);
      const uint32_t a0 = s[0], b0 = s[4], c0 = s[8], d0 = s[12];
      const uint32_t tmp0 = st.add(a0, b0);
      const uint32_t a1 = st.add(tmp0, mx);
      const uint32_t d1 = rotr32(d0 ^ a1, 16);
      const uint32_t c1 = st.add(c0, d1);
      const uint32_t b1 = rotr32(b0 ^ c1, 12);
      const uint32_t tmp1 = st.add(a1, b1);
      const uint32_t a2 = st.add(tmp1, my);
      const uint32_t d2 = rotr32(d1 ^ a2, 8);
      const uint32_t c2 = st.add(c1, d2);
      const uint32_t bn = rotr32(b1 ^ c2, 7);
      st.lin(bn);
      st.lin(d2);
      s[0] = a2;
      s[4] = bn;
      s[8] = c2;
      s[12] = d2;
      if (g == 3) {  // (3,3,3,3) -> (0,1,2,3)
        rot_row(s, 0, 1);
        rot_row(s, 1, 2);
        rot_row(s, 2, 3);
      } else if (g == 7) {  // (3,4,5,6) = (3,0,1,2) -> (0,0,0,0)
        rot_row(s, 0, 1);
        rot_row(s, 2, 3);
        rot_row(s, 3, 2);
      } else {
        rot_row(s, 0, 1);
        rot_row(s, 1, 1);
        rot_row(s, 2, 1);
        rot_row(s, 3, 1);
      }
      sink.checkpoint(st.widx);
    }
  }
  // out_hi = state[w + 8] ^ cv[w] (lin rows), then the zero padding rows.
  for (int w = 0; w < 8; ++w) {
    st.lin(s[8 + w] ^ cv[w]);
    sink.checkpoint(st.widx);
  }
  st.word(st.z, st.a, st.b);  // USEFUL_BITS is not a multiple of 64
  sink.checkpoint(st.widx);
#pragma unroll 1
  while (st.widx < WORDS) {
    st.word(0, 0, 0);
    sink.checkpoint(st.widx);
  }
}

// v_perm_b32 semantics: byte i of the result is byte sel_i of the 8 bytes
// {lo (bytes 0..3), hi (bytes 4..7)}; selectors here are all < 8.
FLK_WHD uint32_t byte_perm(uint32_t hi, uint32_t lo, uint32_t sel) {
#ifdef __HIP_DEVICE_COMPILE__
  return __builtin_amdgcn_perm(hi, lo, sel);
#else
  const uint64_t v = ((uint64_t)hi << 32) | lo;
  uint32_t r = 0;
  for (int i = 0; i < 4; ++i) r |= (uint32_t)((v >> (8 * ((sel >> (8 * i)) & 7))) & 0xff) << (8 * i);
  return r;
#endif
}

// Lincheck stripe chunk of one group of 8 blocks and one word position:
// w[r] = z word of block 8g + r; out (64 bytes as 16 dwords) byte p has bit r
// = bit p of w[r].  Bit butterflies transpose each byte column's 8x8 bit
// matrix, byte butterflies then transpose the 8x8 byte matrix.
FLK_WHD void stripe_chunk(const uint64_t w[8], uint32_t out[16]) {
  uint32_t lo[8], hi[8];
  for (int r = 0; r < 8; ++r) {
    lo[r] = (uint32_t)w[r];
    hi[r] = (uint32_t)(w[r] >> 32);
  }
  const uint32_t masks[3] = {0x0f0f0f0fu, 0x33333333u, 0x55555555u};
  for (int st = 0; st < 3; ++st) {
    const int k = 4 >> st;
    const uint32_t m = masks[st];
    for (int r = 0; r < 8; ++r) {
      if (r & k) continue;
      uint32_t t = ((lo[r] >> k) ^ lo[r + k]) & m;
      lo[r + k] ^= t;
      lo[r] ^= t << k;
      t = ((hi[r] >> k) ^ hi[r + k]) & m;
      hi[r + k] ^= t;
      hi[r] ^= t << k;
    }
  }
  // now word t holds, at byte c, output byte 8c + t
  for (int t = 0; t < 4; ++t) {
    const uint32_t x = hi[t];
    hi[t] = lo[t + 4];
    lo[t + 4] = x;
  }
  for (int t = 0; t < 8; ++t) {
    if (t & 2) continue;
    uint32_t x = lo[t], y = lo[t + 2];
    lo[t] = byte_perm(y, x, 0x05040100u);
    lo[t + 2] = byte_perm(y, x, 0x07060302u);
    x = hi[t];
    y = hi[t + 2];
    hi[t] = byte_perm(y, x, 0x05040100u);
    hi[t + 2] = byte_perm(y, x, 0x07060302u);
  }
  for (int t = 0; t < 8; t += 2) {
    uint32_t x = lo[t], y = lo[t + 1];
    lo[t] = byte_perm(y, x, 0x06020400u);
    lo[t + 1] = byte_perm(y, x, 0x07030501u);
    x = hi[t];
    y = hi[t + 1];
    hi[t] = byte_perm(y, x, 0x06020400u);
    hi[t + 1] = byte_perm(y, x, 0x07030501u);
  }
  for (int c = 0; c < 8; ++c) {
    out[2 * c] = lo[c];
    out[2 * c + 1] = hi[c];
  }
}

}  // namespace wit
}  // namespace flk
