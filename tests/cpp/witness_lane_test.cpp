// Host-side check of the per-compression witness code shared with the device
// kernel (flock_challenge_multi_amd/csrc/witness_lane.hpp) against the CPU
// oracle (oracle/witness.c): the row streams of whole blocks, the order and
// spacing of the sink calls that the kernel's LDS ring relies on, the
// compression pre-pass and the stripe transpose.  Prints "witness lane ok".
#include <cstdio>
#include <cstdlib>
#include <cstring>

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

// Mirrors the kernel's sink: words land in a 20-slot ring per lane and every
// checkpoint writes out at most one 16-word window.
struct HostSink {
  static constexpr int RING = 20, WIN = 16;
  uint64_t z[WORDS], a[WORDS], b[WORDS];
  int next = 0, flushed = 0;
  void put(int widx, uint64_t zw, uint64_t aw, uint64_t bw) {
    CHECK(widx == next, "word %d out of order (expected %d)", widx, next);
    CHECK(widx - flushed < RING, "word %d would overwrite unflushed slot (flushed %d)", widx, flushed);
    CHECK(widx < WORDS, "word %d past the block", widx);
    if (widx < 0 || widx >= WORDS) return;
    z[widx] = zw;
    a[widx] = aw;
    b[widx] = bw;
    next = widx + 1;
  }
  void checkpoint(int widx) {
    if (widx >= flushed + WIN) flushed += WIN;
    CHECK(widx - flushed < WIN, "checkpoint at %d leaves %d pending", widx, widx - flushed);
  }
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
  static constexpr Sched sc = make_sched();
  const uint32_t t_lo = (uint32_t)c.counter, t_hi = (uint32_t)(c.counter >> 32);
  ArrMsg msg{c.m};
  uint32_t s[16];
  compress_state(c.cv, msg, t_lo, t_hi, c.block_len, c.flags, sc, s);
  uint32_t out16[16];
  uint8_t blk[64];
  std::memcpy(blk, c.m, 64);
  orc_blake3_compress(c.cv, blk, c.counter, c.block_len, c.flags, out16);
  uint32_t out_lo[8];
  for (int w = 0; w < 8; ++w) {
    out_lo[w] = s[w] ^ s[w + 8];
    CHECK(out_lo[w] == out16[w], "compress_state out_lo[%d]", w);
  }
  HostSink sink;
  block_rows(c.cv, msg, t_lo, t_hi, c.block_len, c.flags, out_lo, sc, sink);
  CHECK(sink.next == WORDS && sink.flushed == WORDS, "stream ended at %d, flushed %d", sink.next, sink.flushed);
  uint64_t z[WORDS], a[WORDS], b[WORDS];
  orc_blake3_block_witness(&c, z, a, b);
  for (int i = 0; i < WORDS; ++i) {
    CHECK(sink.z[i] == z[i], "z word %d: %016llx vs %016llx", i, (unsigned long long)sink.z[i], (unsigned long long)z[i]);
    CHECK(sink.a[i] == a[i], "a word %d", i);
    CHECK(sink.b[i] == b[i], "b word %d", i);
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
