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

// Message schedule: position j of round r consumes message word PERM^r(j)
// (blake3.rs:149,325-340).
constexpr uint8_t MSG_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
constexpr int sched_idx(int r, int j) {
  for (int i = 0; i < r; ++i) j = MSG_PERM[j];
  return j;
}

FLK_WHD uint32_t iv(int i) {
  return i == 0 ? 0x6a09e667u : i == 1 ? 0xbb67ae85u : i == 2 ? 0x3c6ef372u : i == 3 ? 0xa54ff53au
       : i == 4 ? 0x510e527fu : i == 5 ? 0x9b05688cu : i == 6 ? 0x1f83d9abu : 0x5be0cd19u;
}

FLK_WHD uint32_t rotr32(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// The compression's initial state (blake3.rs:282-300).
FLK_WHD void init_state(uint32_t s[16], const uint32_t cv[8], uint32_t t_lo, uint32_t t_hi, uint32_t blen,
                        uint32_t flags) {
  for (int i = 0; i < 8; ++i) s[i] = cv[i];
  for (int i = 0; i < 4; ++i) s[8 + i] = iv(i);
  s[12] = t_lo;
  s[13] = t_hi;
  s[14] = blen;
  s[15] = flags;
}

// G call G of the compression (textbook lane order, blake3.rs:250-270).
template <int G, class Msg>
FLK_WHD void g_call(uint32_t s[16], const Msg& msg) {
  constexpr int r = G / 8, g = G % 8;
  constexpr int a = g < 4 ? g : g - 4;
  constexpr int b = g < 4 ? 4 + g : 4 + (g - 3) % 4;
  constexpr int c = g < 4 ? 8 + g : 8 + (g - 2) % 4;
  constexpr int d = g < 4 ? 12 + g : 12 + (g - 1) % 4;
  s[a] = s[a] + s[b] + msg(sched_idx(r, 2 * g));
  s[d] = rotr32(s[d] ^ s[a], 16);
  s[c] = s[c] + s[d];
  s[b] = rotr32(s[b] ^ s[c], 12);
  s[a] = s[a] + s[b] + msg(sched_idx(r, 2 * g + 1));
  s[d] = rotr32(s[d] ^ s[a], 8);
  s[c] = s[c] + s[d];
  s[b] = rotr32(s[b] ^ s[c], 7);
}

// All 56 G calls, straight-line (the part that writes words 4..7, out_lo,
// needs the final state before its first word).
template <int G = 0, class Msg>
FLK_WHD void compress_rounds(uint32_t s[16], const Msg& msg) {
  if constexpr (G < 56) {
    g_call<G>(s, msg);
    compress_rounds<G + 1>(s, msg);
  }
}

// State words in caller storage: word i at p[i * stride].
struct StridedState {
  uint32_t* p;
  int stride;
  FLK_WHD uint32_t get(int i) const { return p[i * stride]; }
  FLK_WHD void set(int i, uint32_t v) const { p[i * stride] = v; }
};

// Runs G calls 0 .. n_g - 1 on the state in `st` as one rolled loop (lanes
// and message indices from scalar arithmetic): a part of a block that starts
// after G call k fast-forwards with it.
template <class Msg>
FLK_WHD void advance_state(const StridedState& st, const Msg& msg, int n_g) {
  uint64_t perm = 0;  // MSG_PERM as 16 nibbles
  for (int i = 0; i < 16; ++i) perm |= (uint64_t)MSG_PERM[i] << (4 * i);
#pragma unroll 1
  for (int G = 0; G < n_g; ++G) {
    const int r = G >> 3, g = G & 7, j = g & 3;
    const int a = j;
    const int b = 4 + (g < 4 ? j : (j + 1) & 3);
    const int c = 8 + (g < 4 ? j : (j + 2) & 3);
    const int d = 12 + (g < 4 ? j : (j + 3) & 3);
    int mi = 2 * g, mj = 2 * g + 1;
    for (int i = 0; i < r; ++i) {
      mi = (int)((perm >> (4 * mi)) & 15);
      mj = (int)((perm >> (4 * mj)) & 15);
    }
    uint32_t va = st.get(a), vb = st.get(b), vc = st.get(c), vd = st.get(d);
    va = va + vb + msg(mi);
    vd = rotr32(vd ^ va, 16);
    vc = vc + vd;
    vb = rotr32(vb ^ vc, 12);
    va = va + vb + msg(mj);
    vd = rotr32(vd ^ va, 8);
    vc = vc + vd;
    vb = rotr32(vb ^ vc, 7);
    st.set(a, va);
    st.set(b, vb);
    st.set(c, vc);
    st.set(d, vd);
  }
}

// The three row vectors (z, a, b) of one block, written as u64 words at bit
// positions known at compile time: every push names its first bit.
template <class Sink>
struct Rows {
  Sink& sink;
  uint64_t z, a, b;  // bits of the current word below the next position

  // W bits of each vector (values already masked to W bits) at bit POS.
  template <int POS, int W>
  FLK_WHD void push(uint32_t zv, uint32_t av, uint32_t bv) {
    constexpr int used = POS & 63;
    if constexpr (used == 0) {
      z = zv;
      a = av;
      b = bv;
    } else {
      z |= (uint64_t)zv << used;
      a |= (uint64_t)av << used;
      b |= (uint64_t)bv << used;
    }
    if constexpr (used + W >= 64) {
      sink.put(POS >> 6, z, a, b);
      constexpr int sh = 64 - used;  // bits of this value in the emitted word, 1..W
      z = (uint64_t)zv >> sh;
      a = (uint64_t)av >> sh;
      b = (uint64_t)bv >> sh;
    }
  }
  // A lin row word (blake3.rs:1889-1890): z = a = v, b = 1.
  template <int POS>
  FLK_WHD void lin(uint32_t v) {
    push<POS, 32>(v, v, 0xffffffffu);
  }
  // One 32-bit addition: its 31 carry rows (common.rs:80-88) and the sum.
  template <int POS>
  FLK_WHD uint32_t add(uint32_t x, uint32_t y) {
    const uint32_t sum = x + y;
    const uint32_t cin = sum ^ x ^ y;
    const uint32_t left = (x ^ cin) & 0x7fffffffu;
    const uint32_t right = (y ^ cin) & 0x7fffffffu;
    push<POS, 31>(left & right, left, right);
    return sum;
  }
};

// The row words of one block in 59 steps: step 0 the words before the G
// records, steps 1..56 the G calls, step 57 out_hi, step 58 the zero padding
// words.  run<STEP>() returns the number of words complete after the step; s
// holds the state before the step's G call.
template <class Sink, class Msg>
struct BlockSteps {
  Rows<Sink>& st;
  Sink& sink;
  const Msg& msg;
  const uint32_t* cv;
  uint32_t t_lo, t_hi, blen, flags;
  uint32_t out_lo[8];  // final state[w] ^ state[w + 8], for step 0
  uint32_t s[16];

  template <int STEP>
  FLK_WHD int run() {
    if constexpr (STEP == 0) {
      const uint64_t ones = ~0ull;
      // words 0..3: cv; 4..7: out_lo (aligned lin rows)
      for (int i = 0; i < 4; ++i) {
        const uint64_t v = (uint64_t)cv[2 * i] | ((uint64_t)cv[2 * i + 1] << 32);
        sink.put(i, v, v, ones);
      }
      for (int i = 0; i < 4; ++i) {
        const uint64_t v = (uint64_t)out_lo[2 * i] | ((uint64_t)out_lo[2 * i + 1] << 32);
        sink.put(4 + i, v, v, ones);
      }
      // words 8..17: the constant wire at bit 512, then m[0..16), t_lo, t_hi,
      // block_len, flags from bit 513 on (input rows: z = a = value, b = 1)
      uint32_t prev_top = 1;
      for (int i = 0; i < 10; ++i) {
        const uint32_t lo = 2 * i < 16 ? msg(2 * i) : (2 * i == 16 ? t_lo : blen);
        const uint32_t hi = 2 * i + 1 < 16 ? msg(2 * i + 1) : (2 * i + 1 == 17 ? t_hi : flags);
        const uint64_t v = (uint64_t)prev_top | ((uint64_t)lo << 1) | ((uint64_t)hi << 33);
        sink.put(8 + i, v, v, ones);
        prev_top = hi >> 31;
      }
      st.z = st.a = prev_top;  // flags bit 31 at bit 1152
      st.b = 1;
      return 18;
    } else if constexpr (STEP <= 56) {
      constexpr int G = STEP - 1, r = G / 8, gi = G % 8;
      constexpr int la = gi < 4 ? gi : gi - 4;
      constexpr int lb = gi < 4 ? 4 + gi : 4 + (gi - 3) % 4;
      constexpr int lc = gi < 4 ? 8 + gi : 8 + (gi - 2) % 4;
      constexpr int ld = gi < 4 ? 12 + gi : 12 + (gi - 1) % 4;
      constexpr int base = GS_BASE + G_STRIDE * G;
      const uint32_t mx = msg(sched_idx(r, 2 * gi)), my = msg(sched_idx(r, 2 * gi + 1));
      const uint32_t tmp0 = st.template add<base + 0>(s[la], s[lb]);
      const uint32_t a1 = st.template add<base + 31>(tmp0, mx);
      const uint32_t d1 = rotr32(s[ld] ^ a1, 16);
      const uint32_t c1 = st.template add<base + 62>(s[lc], d1);
      const uint32_t b1 = rotr32(s[lb] ^ c1, 12);
      const uint32_t tmp1 = st.template add<base + 93>(a1, b1);
      const uint32_t a2 = st.template add<base + 124>(tmp1, my);
      const uint32_t d2 = rotr32(d1 ^ a2, 8);
      const uint32_t c2 = st.template add<base + 155>(c1, d2);
      const uint32_t bn = rotr32(b1 ^ c2, 7);
      st.template lin<base + 186>(bn);
      st.template lin<base + 218>(d2);
      s[la] = a2;
      s[lb] = bn;
      s[lc] = c2;
      s[ld] = d2;
      return (base + G_STRIDE) >> 6;
    } else if constexpr (STEP == 57) {
      // out_hi = state[w + 8] ^ cv[w] (lin rows), then the last partial word
#define FLK_OUT_HI(w) st.template lin<OUT_HI_BASE + 32 * (w)>(s[8 + (w)] ^ cv[w])
      FLK_OUT_HI(0);
      FLK_OUT_HI(1);
      FLK_OUT_HI(2);
      FLK_OUT_HI(3);
      FLK_OUT_HI(4);
      FLK_OUT_HI(5);
      FLK_OUT_HI(6);
      FLK_OUT_HI(7);
#undef FLK_OUT_HI
      static_assert(USEFUL_BITS % 64 != 0, "the out_hi rows end inside a word");
      sink.put(USEFUL_BITS >> 6, st.z, st.a, st.b);
      return (USEFUL_BITS >> 6) + 1;
    } else {
      for (int w = (USEFUL_BITS >> 6) + 1; w < WORDS; ++w) sink.put(w, 0, 0, 0);
      return WORDS;
    }
  }
};

constexpr int N_STEPS = 59;

template <int LO, int HI, class Steps>
FLK_WHD int run_step(int step, Steps& f) {
  if constexpr (HI - LO == 1) {
    return f.template run<LO>();
  } else {
    constexpr int MID = (LO + HI) / 2;
    if (step < MID) return run_step<LO, MID>(step, f);
    return run_step<MID, HI>(step, f);
  }
}

// The first step whose words can complete word w_lo - 1 (so that word w_lo
// and later ones are built entirely by the steps that follow).
FLK_WHD int first_step(int w_lo) {
  const int bit = 64 * w_lo - 1;
  if (bit < GS_BASE) return 0;
  if (bit >= OUT_HI_BASE) return 57;
  return 1 + (bit - GS_BASE) / G_STRIDE;
}

// Words [w_lo, w_hi) of one block's (z, a, b).  The sink keeps the words of
// its range, gets a checkpoint after every step (the only flush site; at most
// 19 words pending since the last 16-word boundary it has passed) and answers
// true once its range is written out.  The part with w_lo == 0 first runs the
// whole compression for out_lo (words 4..7); a part that starts after G call
// k fast-forwards the state in the sink's scratch (Sink::state_scratch,
// released by Sink::state_scratch_done).
template <class Sink, class Msg>
FLK_WHD void block_rows(const uint32_t cv[8], const Msg& msg, uint32_t t_lo, uint32_t t_hi, uint32_t blen,
                        uint32_t flags, int w_lo, Sink& sink) {
  Rows<Sink> st{sink, 0, 0, 0};
  BlockSteps<Sink, Msg> f{st, sink, msg, cv, t_lo, t_hi, blen, flags, {}, {}};
  init_state(f.s, cv, t_lo, t_hi, blen, flags);
  const int s0 = first_step(w_lo);
  if (s0 == 0) {
    uint32_t e[16];
    for (int i = 0; i < 16; ++i) e[i] = f.s[i];
    compress_rounds(e, msg);
    for (int i = 0; i < 8; ++i) f.out_lo[i] = e[i] ^ e[i + 8];
  } else if (s0 > 1) {
    const StridedState ss = sink.state_scratch();
    for (int i = 0; i < 16; ++i) ss.set(i, f.s[i]);
    advance_state(ss, msg, s0 - 1);
    for (int i = 0; i < 16; ++i) f.s[i] = ss.get(i);
    sink.state_scratch_done();
  }
#pragma unroll 1
  for (int step = s0; step < N_STEPS; ++step) {
    const int words = run_step<0, N_STEPS>(step, f);
    if (sink.checkpoint(words)) break;
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
