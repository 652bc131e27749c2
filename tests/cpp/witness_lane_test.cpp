// Host-side check of the per-compression witness code shared with the device
// kernel (flock_challenge_multi_amd/csrc/witness_lane.hpp) against the CPU
// oracle (oracle/witness.c): the row words of whole blocks built in 1, 2, 4,
// 8 and 16 parts, the order and spacing of the sink calls that the kernel's
// LDS ring relies on, the compression state and the stripe transpose.  Prints
// "witness lane ok".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>

#include "flock_oracle.h"
#include "witness_lane.hpp"

using namespace flk::wit;

static int fails = 0;
#define CHECK(c, ...)                 \
  do {                                \
    if (!(c)) {                       \
      std::printf(__VA_ARGS__);       \
      std::printf("\n");              \
      if (++fails > 20) std::exit(1); \
    }                                 \
  } while (0)

// Mirrors the kernel's sink for the part [lo, hi) of a block: words land in
// a 20-slot ring and every checkpoint writes out at most one 16-word window.
struct HostSink {
  static constexpr int RING = 20, WIN = 16;
  uint64_t* z;
  uint64_t* a;
  uint64_t* b;
  uint8_t* seen;
  int lo, hi;
  int flushed, last = -1;
  bool done = false;
  void put(int widx, uint64_t zw, uint64_t aw, uint64_t bw) {
    CHECK(widx > last, "word %d after word %d", widx, last);
    if (widx < lo || widx >= hi) {
      // the kernel stores it anyway: its ring slot must hold no pending word
      for (int j = flushed; j <= last && j < hi; ++j)
        CHECK(j < lo || j % RING != widx % RING, "word %d (outside %d..%d) overwrites pending word %d", widx, lo, hi, j);
      last = widx;
      return;
    }
    last = widx;
    CHECK(!done, "word %d after the part ended", widx);
    CHECK(widx - flushed < RING, "word %d would overwrite unflushed slot (flushed %d)", widx, flushed);
    z[widx] = zw;
    a[widx] = aw;
    b[widx] = bw;
    ++seen[widx];
  }
  bool checkpoint(int words) {
    if (words >= flushed + WIN && flushed < hi) flushed += WIN;
    CHECK(flushed >= hi || words - flushed < WIN, "checkpoint at %d leaves %d pending", words, words - flushed);
    done = flushed >= hi;
    return done;
  }
  uint32_t scratch[16];
  StridedState state_scratch() { return StridedState{scratch, 1}; }
  void state_scratch_done() {}
};

struct ArrMsg {
  const uint32_t* m;
  uint32_t operator()(int i) const { return m[i]; }
};

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next_u64() {
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static void check_block(const orc_compression& c) {
  const uint32_t t_lo = (uint32_t)c.counter, t_hi = (uint32_t)(c.counter >> 32);
  ArrMsg msg{c.m};
  uint32_t s[16];
  init_state(s, c.cv, t_lo, t_hi, c.block_len, c.flags);
  advance_state(StridedState{s, 1}, msg, 56);
  uint32_t out16[16];
  uint8_t blk[64];
  std::memcpy(blk, c.m, 64);
  orc_blake3_compress(c.cv, blk, c.counter, c.block_len, c.flags, out16);
  for (int w = 0; w < 8; ++w) CHECK((s[w] ^ s[w + 8]) == out16[w], "advance_state out_lo[%d]", w);
  uint64_t z[WORDS], a[WORDS], b[WORDS];
  orc_blake3_block_witness(&c, z, a, b);
  for (int parts : {1, 2, 4, 8, 16}) {
    uint64_t gz[WORDS] = {}, ga[WORDS] = {}, gb[WORDS] = {};
    uint8_t seen[WORDS] = {};
    for (int p = 0; p < parts; ++p) {
      const int lo = WORDS / parts * p, hi = WORDS / parts * (p + 1);
      HostSink sink{gz, ga, gb, seen, lo, hi, lo};
      block_rows(c.cv, msg, t_lo, t_hi, c.block_len, c.flags, lo, sink);
      CHECK(sink.done, "part %d/%d did not finish", p, parts);
    }
    for (int i = 0; i < WORDS; ++i) {
      CHECK(seen[i] == 1, "parts %d: word %d written %d times", parts, i, seen[i]);
      CHECK(gz[i] == z[i], "parts %d: z word %d: %016llx vs %016llx", parts, i, (unsigned long long)gz[i],
            (unsigned long long)z[i]);
      CHECK(ga[i] == a[i], "parts %d: a word %d", parts, i);
      CHECK(gb[i] == b[i], "parts %d: b word %d", parts, i);
    }
  }
}

int main() {
  orc_compression c;
  std::memset(&c, 0, sizeof c);
  check_block(c);  // the padding compression
  std::memset(&c, 0xff, sizeof c);
  check_block(c);
  for (int t = 0; t < 200; ++t) {
    for (int i = 0; i < 8; ++i) c.cv[i] = (uint32_t)next_u64();
    for (int i = 0; i < 16; ++i) c.m[i] = (uint32_t)next_u64();
    c.counter = next_u64();
    c.block_len = (uint32_t)next_u64();
    c.flags = (uint32_t)next_u64();
    check_block(c);
  }
  for (int t = 0; t < 2000; ++t) {
    uint64_t w[8];
    for (int r = 0; r < 8; ++r) w[r] = (next_u64() & next_u64()) ^ (t & 1 ? ~0ull : 0);
    uint32_t out[16];
    stripe_chunk(w, out);
    const uint8_t* ob = reinterpret_cast<const uint8_t*>(out);
    for (int p = 0; p < 64; ++p) {
      uint8_t want = 0;
      for (int r = 0; r < 8; ++r) want |= (uint8_t)(((w[r] >> p) & 1) << r);
      CHECK(ob[p] == want, "stripe byte %d", p);
    }
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("witness lane ok\n");
  return 0;
}
