// Host-side BLAKE3, written from the BLAKE3 specification: the compression
// function, chunk chaining, PARENT merges and root output blocks read by the
// output counter.  Used by the Fiat-Shamir transcript (transcript.cpp, an
// incremental hasher over a CV stack) and by the R1CS statement digest
// (r1cs_blake3.cpp, a tree hash of a long buffer split over host threads).
// The device kernels have their own copy (blake3_dev.hpp).
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>

namespace flk {
namespace b3 {

constexpr uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                            0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
constexpr int PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
enum : uint32_t { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8 };
constexpr size_t CHUNK_LEN = 1024, BLOCK_LEN = 64;

inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

// Message word order of round r: PERM applied r times to the identity.
struct Schedule {
  uint8_t w[7][16];
  constexpr Schedule() : w() {
    for (int i = 0; i < 16; ++i) w[0][i] = (uint8_t)i;
    for (int r = 1; r < 7; ++r)
      for (int i = 0; i < 16; ++i) w[r][i] = w[r - 1][PERM[i]];
  }
};
constexpr Schedule SCHEDULE{};

// The full 16-word output of the compression function.  The state lives in
// sixteen scalars (not an indexed array) so that it stays in registers, and
// each round reads the message through the precomputed schedule instead of
// permuting a copy.
inline void compress(const uint32_t cv[8], const uint32_t m[16], uint64_t counter, uint32_t block_len,
                     uint32_t flags, uint32_t out[16]) {
  uint32_t s0 = cv[0], s1 = cv[1], s2 = cv[2], s3 = cv[3], s4 = cv[4], s5 = cv[5], s6 = cv[6], s7 = cv[7];
  uint32_t s8 = IV[0], s9 = IV[1], s10 = IV[2], s11 = IV[3];
  uint32_t s12 = (uint32_t)counter, s13 = (uint32_t)(counter >> 32), s14 = block_len, s15 = flags;
  auto g = [](uint32_t& a, uint32_t& b, uint32_t& c, uint32_t& d, uint32_t x, uint32_t y) {
    a = a + b + x;
    d = rotr(d ^ a, 16);
    c = c + d;
    b = rotr(b ^ c, 12);
    a = a + b + y;
    d = rotr(d ^ a, 8);
    c = c + d;
    b = rotr(b ^ c, 7);
  };
#pragma GCC unroll 7
  for (int r = 0; r < 7; ++r) {
    const uint8_t* z = SCHEDULE.w[r];
    g(s0, s4, s8, s12, m[z[0]], m[z[1]]);
    g(s1, s5, s9, s13, m[z[2]], m[z[3]]);
    g(s2, s6, s10, s14, m[z[4]], m[z[5]]);
    g(s3, s7, s11, s15, m[z[6]], m[z[7]]);
    g(s0, s5, s10, s15, m[z[8]], m[z[9]]);
    g(s1, s6, s11, s12, m[z[10]], m[z[11]]);
    g(s2, s7, s8, s13, m[z[12]], m[z[13]]);
    g(s3, s4, s9, s14, m[z[14]], m[z[15]]);
  }
  out[0] = s0 ^ s8, out[1] = s1 ^ s9, out[2] = s2 ^ s10, out[3] = s3 ^ s11;
  out[4] = s4 ^ s12, out[5] = s5 ^ s13, out[6] = s6 ^ s14, out[7] = s7 ^ s15;
  out[8] = s8 ^ cv[0], out[9] = s9 ^ cv[1], out[10] = s10 ^ cv[2], out[11] = s11 ^ cv[3];
  out[12] = s12 ^ cv[4], out[13] = s13 ^ cv[5], out[14] = s14 ^ cv[6], out[15] = s15 ^ cv[7];
}

inline void words_le(const uint8_t* b, uint32_t w[16]) {
  for (int i = 0; i < 16; ++i)
    w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
           ((uint32_t)b[4 * i + 3] << 24);
}

inline void store_words_le(const uint32_t* w, int n, uint8_t* out) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
}

// The last compression of a node, kept unevaluated: its chaining value, or
// (at the root) the output blocks read by counter.
struct Output {
  uint32_t cv[8];
  uint32_t block[16];
  uint64_t counter;
  uint32_t block_len, flags;
  void chaining_value(uint32_t out[8]) const {
    uint32_t o[16];
    compress(cv, block, counter, block_len, flags, o);
    std::memcpy(out, o, 32);
  }
  void root_bytes(uint64_t offset, uint8_t* out, size_t len) const {
    uint64_t blk = offset / 64;
    size_t skip = (size_t)(offset % 64);
    while (len) {
      uint32_t o[16];
      compress(cv, block, blk, block_len, flags | ROOT, o);
      uint8_t bytes[64];
      store_words_le(o, 16, bytes);
      const size_t take = 64 - skip < len ? 64 - skip : len;
      std::memcpy(out, bytes + skip, take);
      out += take;
      len -= take;
      skip = 0;
      ++blk;
    }
  }
};

inline Output parent_output(const uint32_t l[8], const uint32_t r[8]) {
  Output o;
  std::memcpy(o.cv, IV, 32);
  std::memcpy(o.block, l, 32);
  std::memcpy(o.block + 8, r, 32);
  o.counter = 0;
  o.block_len = BLOCK_LEN;
  o.flags = PARENT;
  return o;
}

// The Output of one whole chunk: `len` <= CHUNK_LEN bytes at chunk index `counter`.
inline Output chunk_output(const uint8_t* data, size_t len, uint64_t counter) {
  uint32_t cv[8];
  std::memcpy(cv, IV, 32);
  size_t off = 0;
  uint32_t start = CHUNK_START;
  while (len - off > BLOCK_LEN) {
    uint32_t w[16], o[16];
    words_le(data + off, w);
    compress(cv, w, counter, BLOCK_LEN, start, o);
    std::memcpy(cv, o, 32);
    start = 0;
    off += BLOCK_LEN;
  }
  uint8_t last[64] = {};
  std::memcpy(last, data + off, len - off);
  Output out;
  std::memcpy(out.cv, cv, 32);
  words_le(last, out.block);
  out.counter = counter;
  out.block_len = (uint32_t)(len - off);
  out.flags = start | CHUNK_END;
  return out;
}

// BLAKE3 hash of `len` bytes as a tree of whole chunks split over up to
// `threads` host threads (each subtree of the left-balanced chunk tree is
// independent).  Equal to the sequential hash by construction.
void hash_parallel(const uint8_t* data, size_t len, uint8_t out[32], unsigned threads);

}  // namespace b3
}  // namespace flk
