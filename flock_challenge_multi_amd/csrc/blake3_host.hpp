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

inline void g(uint32_t* s, int a, int b, int c, int d, uint32_t x, uint32_t y) {
  s[a] = s[a] + s[b] + x;
  s[d] = rotr(s[d] ^ s[a], 16);
  s[c] = s[c] + s[d];
  s[b] = rotr(s[b] ^ s[c], 12);
  s[a] = s[a] + s[b] + y;
  s[d] = rotr(s[d] ^ s[a], 8);
  s[c] = s[c] + s[d];
  s[b] = rotr(s[b] ^ s[c], 7);
}

// The full 16-word output of the compression function.
inline void compress(const uint32_t cv[8], const uint32_t m_in[16], uint64_t counter, uint32_t block_len,
                     uint32_t flags, uint32_t out[16]) {
  uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                    IV[0], IV[1], IV[2], IV[3], (uint32_t)counter, (uint32_t)(counter >> 32), block_len, flags};
  uint32_t m[16];
  std::memcpy(m, m_in, sizeof m);
  for (int r = 0; r < 7; ++r) {
    g(s, 0, 4, 8, 12, m[0], m[1]);
    g(s, 1, 5, 9, 13, m[2], m[3]);
    g(s, 2, 6, 10, 14, m[4], m[5]);
    g(s, 3, 7, 11, 15, m[6], m[7]);
    g(s, 0, 5, 10, 15, m[8], m[9]);
    g(s, 1, 6, 11, 12, m[10], m[11]);
    g(s, 2, 7, 8, 13, m[12], m[13]);
    g(s, 3, 4, 9, 14, m[14], m[15]);
    uint32_t p[16];
    for (int i = 0; i < 16; ++i) p[i] = m[PERM[i]];
    std::memcpy(m, p, sizeof m);
  }
  for (int i = 0; i < 8; ++i) {
    out[i] = s[i] ^ s[i + 8];
    out[i + 8] = s[i + 8] ^ cv[i];
  }
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
