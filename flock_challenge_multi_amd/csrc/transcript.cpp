// Fiat–Shamir transcript over BLAKE3, host side (CF = crates/flock-core/src/):
//
//   flk_fs_*   FsChallenger::with_hash(domain, HashKind::Blake3) and its
//              Challenger methods, CF/challenger.rs:170-177,226-383,384-474,502-517
//
// The transcript is an incremental BLAKE3 hasher over a byte stream:
//   new(domain)            0x01 || u64le(len) || domain
//   observe_label(l)       0x02 || u64le(len) || l
//   observe_f128(v)        0x03 0x01 || lo_le || hi_le
//   observe_f128_slice(v)  0x03 0x02 || u64le(n) || (lo_le || hi_le)*
//   observe_bytes(b)       0x05 || u64le(len) || b
//   sample_f128()          absorb 0x04 0x01, read 16 bytes of the XOF of the
//                          state so far (the state is not finalised), absorb them
//   sample_f128_vec(n)     absorb 0x04 0x02 || u64le(n), read 16 n XOF bytes,
//                          absorb them
//   state_digest()         the 32-byte BLAKE3 hash of the stream so far
//   sample_distinct_queries(block_len, count)   sample_f128 until `count`
//                          distinct q = lo % block_len (first occurrence kept),
//                          sorted (CF/pcs/ligerito.rs:9665-9720)
//   grind_pow(bits)        nonce = flk_pow_grind(state_digest, bits) (device),
//                          then observe_bytes(u64le(nonce))
//   verify_pow(n, bits)    the PoW check against state_digest (n == 0 when
//                          bits == 0), then observe_bytes(u64le(n))
//
// BLAKE3 comes from blake3_host.hpp (the specification's compression, chunk
// chaining and root output blocks); this file keeps the incremental CV stack.  The SHA-256
// transcript of the reference is not provided: the prover's BLAKE3 profile
// uses this one.
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/flock_gpu.h"
#include "blake3_host.hpp"

namespace flk {
int set_error(int code, const char* fmt, ...);
int check_enabled();
}  // namespace flk

namespace {

using flk::b3::BLOCK_LEN;
using flk::b3::CHUNK_END;
using flk::b3::CHUNK_LEN;
using flk::b3::CHUNK_START;
using flk::b3::compress;
using flk::b3::IV;
using flk::b3::Output;
using flk::b3::parent_output;
using flk::b3::words_le;

struct ChunkState {
  uint32_t cv[8];
  uint64_t counter = 0;
  uint8_t block[64] = {};
  uint32_t block_len = 0, blocks_done = 0;
  void reset(const uint32_t key[8], uint64_t ctr) {
    std::memcpy(cv, key, 32);
    counter = ctr;
    std::memset(block, 0, sizeof block);
    block_len = 0;
    blocks_done = 0;
  }
  size_t len() const { return (size_t)blocks_done * BLOCK_LEN + block_len; }
  uint32_t start_flag() const { return blocks_done == 0 ? CHUNK_START : 0; }
  void update(const uint8_t* in, size_t n) {
    while (n) {
      if (block_len == BLOCK_LEN) {  // a full block with more input behind it
        uint32_t w[16], o[16];
        words_le(block, w);
        compress(cv, w, counter, BLOCK_LEN, start_flag(), o);
        std::memcpy(cv, o, 32);
        ++blocks_done;
        std::memset(block, 0, sizeof block);
        block_len = 0;
      }
      const size_t take = BLOCK_LEN - block_len < n ? BLOCK_LEN - block_len : n;
      std::memcpy(block + block_len, in, take);
      block_len += (uint32_t)take;
      in += take;
      n -= take;
    }
  }
  Output output() const {
    Output o;
    std::memcpy(o.cv, cv, 32);
    words_le(block, o.block);
    o.counter = counter;
    o.block_len = block_len;
    o.flags = start_flag() | CHUNK_END;
    return o;
  }
};

struct Hasher {
  ChunkState chunk;
  uint32_t stack[64][8];
  int depth = 0;
  Hasher() { chunk.reset(IV, 0); }
  void push_chunk_cv(uint32_t cv[8], uint64_t total_chunks) {
    // merge completed subtrees: one merge per trailing zero of the chunk count
    while ((total_chunks & 1) == 0) {
      uint32_t merged[8];
      parent_output(stack[--depth], cv).chaining_value(merged);
      std::memcpy(cv, merged, 32);
      total_chunks >>= 1;
    }
    std::memcpy(stack[depth++], cv, 32);
  }
  void update(const uint8_t* in, size_t n) {
    while (n) {
      if (chunk.len() == CHUNK_LEN) {  // a full chunk with more input behind it
        uint32_t cv[8];
        chunk.output().chaining_value(cv);
        const uint64_t total = chunk.counter + 1;
        push_chunk_cv(cv, total);
        chunk.reset(IV, total);
      }
      const size_t room = CHUNK_LEN - chunk.len();
      const size_t take = room < n ? room : n;
      chunk.update(in, take);
      in += take;
      n -= take;
    }
  }
  Output root() const {
    Output o = chunk.output();
    for (int i = depth - 1; i >= 0; --i) {
      uint32_t cv[8];
      o.chaining_value(cv);
      o = parent_output(stack[i], cv);
    }
    return o;
  }
};

}  // namespace

struct flk_fs {
  Hasher h;
  void absorb(const void* p, size_t n) { h.update(static_cast<const uint8_t*>(p), n); }
  void absorb_u64(uint64_t v) {
    uint8_t b[8];
    for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(v >> (8 * i));
    absorb(b, 8);
  }
  void absorb_f128(flk_f128 v) {
    absorb_u64(v.lo);
    absorb_u64(v.hi);
  }
  void digest(uint8_t out[32]) const { h.root().root_bytes(0, out, 32); }
  void squeeze(uint8_t* out, size_t n) const { h.root().root_bytes(0, out, n); }
};

using flk::check_enabled;
using flk::set_error;

namespace {
bool pow_hash_ok(const uint8_t digest[32], uint64_t nonce, uint32_t bits) {
  uint8_t pre[64] = {};
  std::memcpy(pre, digest, 32);
  for (int i = 0; i < 8; ++i) pre[32 + i] = (uint8_t)(nonce >> (8 * i));
  Hasher hh;
  hh.update(pre, 64);
  uint8_t h[32];
  hh.root().root_bytes(0, h, 32);
  const uint32_t full = bits / 8, extra = bits % 8;
  for (uint32_t i = 0; i < full; ++i)
    if (h[i]) return false;
  return extra == 0 || (h[full] >> (8 - extra)) == 0;
}
}  // namespace

extern "C" {

int flk_fs_new(const uint8_t* domain, size_t len, flk_fs** out) {
  if (!out || (len && !domain)) return set_error(FLK_E_SHAPE, "fs_new: null pointer");
  flk_fs* t = new (std::nothrow) flk_fs();
  if (!t) return set_error(FLK_E_NOMEM, "fs_new: out of host memory");
  const uint8_t op = 0x01;
  t->absorb(&op, 1);
  t->absorb_u64(len);
  if (len) t->absorb(domain, len);
  *out = t;
  return FLK_OK;
}

int flk_fs_clone(const flk_fs* t, flk_fs** out) {
  if (!t || !out) return set_error(FLK_E_SHAPE, "fs_clone: null pointer");
  flk_fs* c = new (std::nothrow) flk_fs(*t);
  if (!c) return set_error(FLK_E_NOMEM, "fs_clone: out of host memory");
  *out = c;
  return FLK_OK;
}

void flk_fs_free(flk_fs* t) { delete t; }

int flk_fs_observe_label(flk_fs* t, const uint8_t* label, size_t len) {
  if (!t || (len && !label)) return set_error(FLK_E_SHAPE, "fs_observe_label: null pointer");
  const uint8_t op = 0x02;
  t->absorb(&op, 1);
  t->absorb_u64(len);
  if (len) t->absorb(label, len);
  return FLK_OK;
}

int flk_fs_observe_f128(flk_fs* t, flk_f128 v) {
  if (!t) return set_error(FLK_E_SHAPE, "fs_observe_f128: null pointer");
  const uint8_t op[2] = {0x03, 0x01};
  t->absorb(op, 2);
  t->absorb_f128(v);
  return FLK_OK;
}

int flk_fs_observe_f128_slice(flk_fs* t, const flk_f128* v, size_t n) {
  if (!t || (n && !v)) return set_error(FLK_E_SHAPE, "fs_observe_f128_slice: null pointer");
  const uint8_t op[2] = {0x03, 0x02};
  t->absorb(op, 2);
  t->absorb_u64(n);
  for (size_t i = 0; i < n; ++i) t->absorb_f128(v[i]);
  return FLK_OK;
}

int flk_fs_observe_bytes(flk_fs* t, const uint8_t* bytes, size_t len) {
  if (!t || (len && !bytes)) return set_error(FLK_E_SHAPE, "fs_observe_bytes: null pointer");
  const uint8_t op = 0x05;
  t->absorb(&op, 1);
  t->absorb_u64(len);
  if (len) t->absorb(bytes, len);
  return FLK_OK;
}

int flk_fs_sample_f128_vec(flk_fs* t, size_t n, flk_f128* out) {
  if (!t || (n && !out)) return set_error(FLK_E_SHAPE, "fs_sample_f128_vec: null pointer");
  const uint8_t op[2] = {0x04, 0x02};
  t->absorb(op, 2);
  t->absorb_u64(n);
  std::vector<uint8_t> buf(16 * n);
  t->squeeze(buf.data(), buf.size());
  t->absorb(buf.data(), buf.size());
  for (size_t i = 0; i < n; ++i) {
    uint64_t lo = 0, hi = 0;
    for (int j = 0; j < 8; ++j) {
      lo |= (uint64_t)buf[16 * i + j] << (8 * j);
      hi |= (uint64_t)buf[16 * i + 8 + j] << (8 * j);
    }
    out[i] = flk_f128{lo, hi};
  }
  return FLK_OK;
}

int flk_fs_sample_f128(flk_fs* t, flk_f128* out) {
  if (!t || !out) return set_error(FLK_E_SHAPE, "fs_sample_f128: null pointer");
  const uint8_t op[2] = {0x04, 0x01};
  t->absorb(op, 2);
  uint8_t buf[16];
  t->squeeze(buf, 16);
  t->absorb(buf, 16);
  uint64_t lo = 0, hi = 0;
  for (int j = 0; j < 8; ++j) {
    lo |= (uint64_t)buf[j] << (8 * j);
    hi |= (uint64_t)buf[8 + j] << (8 * j);
  }
  *out = flk_f128{lo, hi};
  return FLK_OK;
}

int flk_fs_sample_distinct_queries(flk_fs* t, uint64_t block_len, size_t count, uint64_t* out) {
  if (!t || (count && !out)) return set_error(FLK_E_SHAPE, "fs_sample_distinct_queries: null pointer");
  if (count > block_len)
    return set_error(FLK_E_SHAPE, "fs_sample_distinct_queries: count %zu > block_len %llu", count,
                     (unsigned long long)block_len);
  std::vector<uint64_t> got;
  got.reserve(count);
  while (got.size() < count) {
    flk_f128 v;
    flk_fs_sample_f128(t, &v);
    const uint64_t q = v.lo % block_len;
    bool seen = false;
    for (uint64_t x : got) seen = seen || x == q;
    if (!seen) got.push_back(q);
  }
  std::sort(got.begin(), got.end());
  std::copy(got.begin(), got.end(), out);
  return FLK_OK;
}

int flk_fs_state_digest(const flk_fs* t, uint8_t out[32]) {
  if (!t || !out) return set_error(FLK_E_SHAPE, "fs_state_digest: null pointer");
  t->digest(out);
  return FLK_OK;
}

int flk_fs_grind_pow(flk_fs* t, uint32_t bits, uint64_t* nonce_out, void* stream) {
  if (!t || !nonce_out) return set_error(FLK_E_SHAPE, "fs_grind_pow: null pointer");
  uint64_t nonce = 0;
  if (bits) {
    uint8_t d[32];
    t->digest(d);
    const int rc = flk_pow_grind(d, bits, &nonce, stream);
    if (rc) return rc;  // nothing absorbed: the transcript is unchanged
  }
  uint8_t le[8];
  for (int i = 0; i < 8; ++i) le[i] = (uint8_t)(nonce >> (8 * i));
  flk_fs_observe_bytes(t, le, 8);
  *nonce_out = nonce;
  return FLK_OK;
}

int flk_fs_verify_pow(flk_fs* t, uint64_t nonce, uint32_t bits, int* ok_out) {
  if (!t || !ok_out) return set_error(FLK_E_SHAPE, "fs_verify_pow: null pointer");
  if (bits > 256) return set_error(FLK_E_SHAPE, "fs_verify_pow: bits=%u > 256", bits);
  uint8_t d[32];
  t->digest(d);
  const bool ok = bits == 0 ? nonce == 0 : pow_hash_ok(d, nonce, bits);
  uint8_t le[8];
  for (int i = 0; i < 8; ++i) le[i] = (uint8_t)(nonce >> (8 * i));
  flk_fs_observe_bytes(t, le, 8);  // absorbed whatever the outcome (CF/challenger.rs:518-522)
  *ok_out = ok ? 1 : 0;
  return FLK_OK;
}

}  // extern "C"
