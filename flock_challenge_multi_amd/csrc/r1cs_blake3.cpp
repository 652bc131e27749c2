// The BLAKE3 compression R1CS, host side (CP = crates/flock-prover/src/,
// CF = crates/flock-core/src/).
//
// build_matrices (CP/r1cs_hashes/blake3.rs:483-632) tracks every intermediate
// bit of the compression as a GF(2) combination of witness slots ("Option D":
// sum bits are never materialised, a carry row i of x + y is
// (x_i + cin_i) * (y_i + cin_i) with cin_i = XOR of the carry slots below i).
// Its rows are slot lists passed through xor_dedup (sort, cancel pairs), i.e.
// GF(2) sets.  Here every symbolic bit is a K-bit set, so XOR is a word-wise
// XOR and the row's ascending column list is its set bits in order.
//
// The statement digest (CF/r1cs.rs:334-385) hashes ~21 M column indices; the
// byte stream is built once per m and tree-hashed over host threads
// (blake3_host.hpp), then cached.
#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "blake3_host.hpp"
#include "r1cs.hpp"

namespace flk {
namespace b3 {

namespace {
void subtree_cv(const uint8_t* data, size_t len, uint64_t chunk0, uint32_t out[8], unsigned threads) {
  const size_t n = (len + CHUNK_LEN - 1) / CHUNK_LEN;
  if (n <= 1) {
    chunk_output(data, len, chunk0).chaining_value(out);
    return;
  }
  size_t left = 1;  // the largest power of two number of chunks that leaves one for the right
  while (2 * left < n) left *= 2;
  const size_t left_len = left * CHUNK_LEN;
  uint32_t l[8], r[8];
  if (threads > 1) {
    std::thread t([&] { subtree_cv(data, left_len, chunk0, l, threads / 2); });
    subtree_cv(data + left_len, len - left_len, chunk0 + left, r, threads - threads / 2);
    t.join();
  } else {
    subtree_cv(data, left_len, chunk0, l, 1);
    subtree_cv(data + left_len, len - left_len, chunk0 + left, r, 1);
  }
  parent_output(l, r).chaining_value(out);
}
}  // namespace

void hash_parallel(const uint8_t* data, size_t len, uint8_t out[32], unsigned threads) {
  const size_t n = (len + CHUNK_LEN - 1) / CHUNK_LEN;
  if (n <= 1) {
    chunk_output(data, len, 0).root_bytes(0, out, 32);
    return;
  }
  size_t left = 1;
  while (2 * left < n) left *= 2;
  const size_t left_len = left * CHUNK_LEN;
  uint32_t l[8], r[8];
  if (threads < 2) threads = 2;
  std::thread t([&] { subtree_cv(data, left_len, 0, l, threads / 2); });
  subtree_cv(data + left_len, len - left_len, left, r, threads - threads / 2);
  t.join();
  parent_output(l, r).root_bytes(0, out, 32);
}

}  // namespace b3

namespace r1cs {
namespace {

constexpr size_t W = K / 64;  // u64 words per K-bit set
constexpr size_t WORD_BITS = 32;

// Slot positions (blake3.rs:189-241).
constexpr size_t CV_BASE = 0, OUT_LO_BASE = 256, M_BASE = 513, T_LO_BASE = 1025, T_HI_BASE = 1057,
                 BLEN_BASE = 1089, FLAGS_BASE = 1121, GS_BASE = 1153, G_STRIDE = 250, N_G = 56,
                 OUT_HI_BASE = GS_BASE + N_G * G_STRIDE;
static_assert(OUT_HI_BASE + 8 * WORD_BITS == USEFUL_BITS, "layout");
constexpr size_t carry_slot(size_t g, size_t add) { return GS_BASE + G_STRIDE * g + 31 * add; }
constexpr size_t lin_slot(size_t g, size_t which) { return GS_BASE + G_STRIDE * g + 6 * 31 + 32 * which; }

constexpr uint32_t IV[8] = {0x6a09e667u, 0xbb67ae85u, 0x3c6ef372u, 0xa54ff53au,
                            0x510e527fu, 0x9b05688cu, 0x1f83d9abu, 0x5be0cd19u};
constexpr int MSG_PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};
// G lanes (a, b, c, d): four column then four diagonal calls; message pair g
// of a round reads permuted words (2g, 2g+1) (blake3.rs:153-184).
constexpr int G_LANES[8][4] = {{0, 4, 8, 12}, {1, 5, 9, 13}, {2, 6, 10, 14}, {3, 7, 11, 15},
                               {0, 5, 10, 15}, {1, 6, 11, 12}, {2, 7, 8, 13}, {3, 4, 9, 14}};

inline void set_bit(uint64_t* s, size_t p) { s[p >> 6] ^= uint64_t(1) << (p & 63); }
inline void xor_into(uint64_t* d, const uint64_t* a) {
  for (size_t i = 0; i < W; ++i) d[i] ^= a[i];
}

// A 32-bit symbolic word: bit i is a K-bit set of slots whose XOR is that bit.
struct Sym {
  std::vector<uint64_t> v = std::vector<uint64_t>(WORD_BITS * W, 0);
  uint64_t* bit(size_t i) { return v.data() + i * W; }
  const uint64_t* bit(size_t i) const { return v.data() + i * W; }
  static Sym slots(size_t base) {
    Sym s;
    for (size_t i = 0; i < WORD_BITS; ++i) set_bit(s.bit(i), base + i);
    return s;
  }
  static Sym constant(uint32_t val) {
    Sym s;
    for (size_t i = 0; i < WORD_BITS; ++i)
      if ((val >> i) & 1u) set_bit(s.bit(i), CONST_POS);
    return s;
  }
};

Sym sym_xor(const Sym& x, const Sym& y) {
  Sym o = x;
  for (size_t i = 0; i < o.v.size(); ++i) o.v[i] ^= y.v[i];
  return o;
}
Sym sym_rotr(const Sym& x, size_t n) {
  Sym o;
  for (size_t i = 0; i < WORD_BITS; ++i) std::copy_n(x.bit((i + n) % WORD_BITS), W, o.bit(i));
  return o;
}

struct Builder {
  std::vector<uint64_t> a = std::vector<uint64_t>(K * W, 0), b = std::vector<uint64_t>(K * W, 0);
  uint64_t* arow(size_t r) { return a.data() + r * W; }
  uint64_t* brow(size_t r) { return b.data() + r * W; }

  // Row s: A = {s}, B = {const}: an input or constant wire.
  void input_row(size_t s) {
    set_bit(arow(s), s);
    set_bit(brow(s), CONST_POS);
  }
  // Rows [base, base + 32): A = word bit i, B = {const} (a materialised linear word).
  void lin_rows(size_t base, const Sym& w) {
    for (size_t i = 0; i < WORD_BITS; ++i) {
      std::copy_n(w.bit(i), W, arow(base + i));
      set_bit(brow(base + i), CONST_POS);
    }
  }
  // x + y with carry slots [cb, cb + 31): the 31 carry rows, and the sum word
  // (write_add_carry_rows + Word::add_sum, blake3.rs:418-440,403-413).
  Sym add(const Sym& x, const Sym& y, size_t cb) {
    Sym sum;
    std::vector<uint64_t> cin(W, 0);
    for (size_t i = 0; i < WORD_BITS; ++i) {
      if (i < 31) {
        uint64_t* ar = arow(cb + i);
        uint64_t* br = brow(cb + i);
        std::copy_n(x.bit(i), W, ar);
        xor_into(ar, cin.data());
        std::copy_n(y.bit(i), W, br);
        xor_into(br, cin.data());
      }
      uint64_t* s = sum.bit(i);
      std::copy_n(x.bit(i), W, s);
      xor_into(s, y.bit(i));
      xor_into(s, cin.data());
      if (i < 31) set_bit(cin.data(), cb + i);
    }
    return sum;
  }
};

Matrices build() {
  Builder bl;
  set_bit(bl.arow(CONST_POS), CONST_POS);
  set_bit(bl.brow(CONST_POS), CONST_POS);
  for (size_t j = 0; j < 8 * WORD_BITS; ++j) bl.input_row(CV_BASE + j);
  for (size_t j = 0; j < 16 * WORD_BITS; ++j) bl.input_row(M_BASE + j);
  for (size_t base : {T_LO_BASE, T_HI_BASE, BLEN_BASE, FLAGS_BASE})
    for (size_t j = 0; j < WORD_BITS; ++j) bl.input_row(base + j);

  std::array<Sym, 16> st;
  for (size_t w = 0; w < 8; ++w) st[w] = Sym::slots(CV_BASE + WORD_BITS * w);
  for (size_t i = 0; i < 4; ++i) st[8 + i] = Sym::constant(IV[i]);
  st[12] = Sym::slots(T_LO_BASE);
  st[13] = Sym::slots(T_HI_BASE);
  st[14] = Sym::slots(BLEN_BASE);
  st[15] = Sym::slots(FLAGS_BASE);

  int perm[16];
  for (int i = 0; i < 16; ++i) perm[i] = i;
  for (size_t r = 0; r < 7; ++r) {
    for (size_t gi = 0; gi < 8; ++gi) {
      const size_t g = 8 * r + gi;
      const int la = G_LANES[gi][0], lb = G_LANES[gi][1], lc = G_LANES[gi][2], ld = G_LANES[gi][3];
      const Sym mx = Sym::slots(M_BASE + WORD_BITS * perm[2 * gi]);
      const Sym my = Sym::slots(M_BASE + WORD_BITS * perm[2 * gi + 1]);
      const Sym tmp0 = bl.add(st[la], st[lb], carry_slot(g, 0));
      const Sym a1 = bl.add(tmp0, mx, carry_slot(g, 1));
      const Sym d1 = sym_rotr(sym_xor(st[ld], a1), 16);
      const Sym c1 = bl.add(st[lc], d1, carry_slot(g, 2));
      const Sym b1 = sym_rotr(sym_xor(st[lb], c1), 12);
      const Sym tmp1 = bl.add(a1, b1, carry_slot(g, 3));
      const Sym a2 = bl.add(tmp1, my, carry_slot(g, 4));
      const Sym d2 = sym_rotr(sym_xor(d1, a2), 8);
      const Sym c2 = bl.add(c1, d2, carry_slot(g, 5));
      const Sym b_new = sym_rotr(sym_xor(b1, c2), 7);
      bl.lin_rows(lin_slot(g, 0), b_new);
      bl.lin_rows(lin_slot(g, 1), d2);
      st[la] = a2;
      st[lb] = Sym::slots(lin_slot(g, 0));
      st[lc] = c2;
      st[ld] = Sym::slots(lin_slot(g, 1));
    }
    int next[16];
    for (int i = 0; i < 16; ++i) next[i] = perm[MSG_PERM[i]];
    std::copy(next, next + 16, perm);
  }
  for (size_t w = 0; w < 8; ++w) {
    bl.lin_rows(OUT_LO_BASE + WORD_BITS * w, sym_xor(st[w], st[w + 8]));
    bl.lin_rows(OUT_HI_BASE + WORD_BITS * w, sym_xor(st[w + 8], Sym::slots(CV_BASE + WORD_BITS * w)));
  }

  Matrices m;
  auto to_csr = [](const std::vector<uint64_t>& rows, std::vector<uint32_t>& ptr, std::vector<uint32_t>& cols) {
    ptr.assign(K + 1, 0);
    for (size_t r = 0; r < K; ++r) {
      const uint64_t* s = rows.data() + r * W;
      for (size_t w = 0; w < W; ++w)
        for (uint64_t x = s[w]; x; x &= x - 1) cols.push_back(uint32_t(64 * w + __builtin_ctzll(x)));
      ptr[r + 1] = uint32_t(cols.size());
    }
  };
  to_csr(bl.a, m.a_ptr, m.a_cols);
  to_csr(bl.b, m.b_ptr, m.b_cols);
  return m;
}

Csc transpose(const Matrices& m) {
  Csc c;
  auto one = [](const std::vector<uint32_t>& ptr, const std::vector<uint32_t>& cols, std::vector<uint32_t>& cptr,
                std::vector<uint16_t>& rows) {
    cptr.assign(K + 1, 0);
    for (uint32_t col : cols) ++cptr[col + 1];
    for (size_t i = 0; i < K; ++i) cptr[i + 1] += cptr[i];
    std::vector<uint32_t> fill(cptr.begin(), cptr.end() - 1);
    rows.resize(cols.size());
    for (size_t r = 0; r < K; ++r)
      for (uint32_t j = ptr[r]; j < ptr[r + 1]; ++j) rows[fill[cols[j]]++] = uint16_t(r);
  };
  one(m.a_ptr, m.a_cols, c.a_ptr, c.a_rows);
  one(m.b_ptr, m.b_cols, c.b_ptr, c.b_rows);
  return c;
}

void put_u64(std::vector<uint8_t>& o, uint64_t v) {
  for (int i = 0; i < 8; ++i) o.push_back(uint8_t(v >> (8 * i)));
}

void absorb_matrix(std::vector<uint8_t>& o, const std::vector<uint32_t>& ptr, const std::vector<uint32_t>& cols) {
  put_u64(o, K);
  put_u64(o, K);
  for (size_t r = 0; r < K; ++r) {
    put_u64(o, ptr[r + 1] - ptr[r]);
    for (uint32_t j = ptr[r]; j < ptr[r + 1]; ++j) put_u64(o, cols[j]);
  }
}

}  // namespace

const Matrices& blake3_matrices() {
  static const Matrices m = build();
  return m;
}

const Csc& blake3_csc() {
  static const Csc c = transpose(blake3_matrices());
  return c;
}

void statement_digest(size_t m, uint8_t out[32]) {
  static std::mutex mu;
  static std::map<size_t, std::array<uint8_t, 32>> cache;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(m);
    if (it != cache.end()) {
      std::copy(it->second.begin(), it->second.end(), out);
      return;
    }
  }
  const Matrices& mt = blake3_matrices();
  std::vector<uint8_t> s;
  s.reserve(64 + 8 * (3 * K + 6 + mt.a_cols.size() + mt.b_cols.size() + K));
  const char* label = "flock-r1cs-stmt-v1";
  s.insert(s.end(), label, label + 18);
  put_u64(s, m);
  put_u64(s, K_LOG);
  put_u64(s, K_SKIP);
  s.push_back(0);  // WitnessLayout::RowMajor
  absorb_matrix(s, mt.a_ptr, mt.a_cols);
  absorb_matrix(s, mt.b_ptr, mt.b_cols);
  put_u64(s, K);  // C_0 = I_K
  put_u64(s, K);
  for (size_t r = 0; r < K; ++r) {
    put_u64(s, 1);
    put_u64(s, r);
  }
  unsigned th = std::thread::hardware_concurrency();
  th = th < 2 ? 2 : (th > 16 ? 16 : th);
  std::array<uint8_t, 32> d;
  b3::hash_parallel(s.data(), s.size(), d.data(), th);
  std::lock_guard<std::mutex> g(mu);
  cache[m] = d;
  std::copy(d.begin(), d.end(), out);
}

}  // namespace r1cs
}  // namespace flk
