// flock_host.hpp — C++ mirror of the reference's Rust hot-path interface over the
// C ABI of include/flock_gpu.h.  Names and argument meaning follow the Rust
// entry points (`CF/` = crates/flock-core/src/); device-resident data lives in
// DeviceVec<T> (RAII over flk_alloc/flk_free).  A non-OK status throws GpuError
// carrying the library message; shape errors are raised before any launch.
#pragma once

#include <array>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "flock_gpu.h"

namespace flock {

using F128 = flk_f128;  // #[repr(C, align(16))] { lo: u64, hi: u64 } (CF/field/gf2_128.rs:22-27)
using Hash = std::array<uint8_t, 32>;

struct GpuError : std::runtime_error {
  int code;
  GpuError(int c, const std::string& where) : std::runtime_error(where + ": " + flk_last_error()), code(c) {}
};

inline void check(int rc, const char* where) {
  if (rc != FLK_OK) throw GpuError(rc, where);
}

inline F128 operator+(F128 a, F128 b) { return F128{a.lo ^ b.lo, a.hi ^ b.hi}; }
inline F128 operator*(F128 a, F128 b) { return flk_f128_mul(a, b); }
inline bool operator==(F128 a, F128 b) { return a.lo == b.lo && a.hi == b.hi; }

// Device buffer of n elements of T in stream-ordered pool memory of one
// stream: the allocation, uploads, read-backs and the free are all ordered on
// that stream, so work the caller enqueues on the same stream needs no further
// synchronisation.  Work on another stream that touches the buffer must be
// ordered against this one by the caller.  Uploads are staged by
// flk_copy_to_device (the library's pinned upload ring), so the host vector may
// be released as soon as they return and nothing waits; read-backs wait for the
// stream.
template <class T>
class DeviceVec {
 public:
  DeviceVec() = default;
  explicit DeviceVec(size_t n, void* stream = nullptr) : n_(n), stream_(stream) {
    void* p = nullptr;
    check(flk_alloc_async(n * sizeof(T), stream_, &p), "flk_alloc_async");
    p_ = static_cast<T*>(p);
  }
  explicit DeviceVec(const std::vector<T>& host, void* stream = nullptr) : DeviceVec(host.size(), stream) {
    upload(host);
  }
  ~DeviceVec() {
    if (p_) flk_free_async(p_, stream_);
  }
  DeviceVec(DeviceVec&& o) noexcept : p_(o.p_), n_(o.n_), stream_(o.stream_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceVec& operator=(DeviceVec&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    std::swap(stream_, o.stream_);
    return *this;
  }
  DeviceVec(const DeviceVec&) = delete;
  DeviceVec& operator=(const DeviceVec&) = delete;

  T* data() { return p_; }
  const T* data() const { return p_; }
  size_t len() const { return n_; }
  void* stream() const { return stream_; }

  void upload(const std::vector<T>& host) {
    if (host.size() != n_) throw std::invalid_argument("DeviceVec::upload: length mismatch");
    if (!n_) return;
    check(flk_copy_to_device(p_, host.data(), n_ * sizeof(T), stream_), "flk_copy_to_device");
  }
  // The last element (e.g. a Merkle tree's root) without copying the rest.
  T back() const {
    if (!n_) throw std::out_of_range("DeviceVec::back: empty");
    T h;
    check(flk_read_to_host(&h, p_ + (n_ - 1), sizeof(T), stream_), "flk_read_to_host");
    return h;
  }
  std::vector<T> to_host() const {
    std::vector<T> h(n_);
    if (n_) check(flk_read_to_host(h.data(), p_, n_ * sizeof(T), stream_), "flk_read_to_host");
    else check(flk_synchronize(stream_), "flk_synchronize");
    return h;
  }

 private:
  T* p_ = nullptr;
  size_t n_ = 0;
  void* stream_ = nullptr;
};

// The n results of the reduction that `launch` enqueues (a round message), read
// back through an armed read-back (flk_read_arm / flk_read_armed): the
// reduction's last work-group stores them straight into host memory, so no sum
// or publish kernel runs between the reduction and the host.
template <class T, class Launch>
inline std::vector<T> read_reduction(const DeviceVec<T>& out, size_t n, void* stream, Launch&& launch) {
  std::vector<T> h(n);
  check(flk_read_arm(h.data(), n * sizeof(T)), "flk_read_arm");
  launch();
  check(flk_read_armed(out.data(), n * sizeof(T), stream), "flk_read_armed");
  return h;
}

inline bool is_pow2(size_t n) { return n && !(n & (n - 1)); }
inline size_t log2_exact(size_t n) {
  if (!is_pow2(n)) throw std::invalid_argument("length must be a power of two");
  size_t l = 0;
  while ((size_t(1) << l) < n) ++l;
  return l;
}

// ---------------------------------------------------------------------------
// crate::ntt::AdditiveNttF128 (CF/ntt/additive_ntt_f128.rs:2661-2830)
// ---------------------------------------------------------------------------
namespace ntt {
class AdditiveNttF128 {
 public:
  static AdditiveNttF128 standard(size_t dim) { return AdditiveNttF128(dim); }
  size_t log_domain_size() const { return dim_; }
  F128 twiddle(size_t layer, size_t block) const {
    if (layer >= dim_ || block >= (size_t(1) << layer)) throw std::out_of_range("twiddle");
    if (tw_.empty()) {
      tw_.resize((size_t(1) << dim_) - 1);
      check(flk_ntt_twiddles(dim_, tw_.data()), "flk_ntt_twiddles");
    }
    return tw_[(size_t(1) << layer) - 1 + block];
  }
  void rs_encode_interleaved(const DeviceVec<F128>& msg, DeviceVec<F128>& codeword, size_t num_ntts,
                             void* stream = nullptr) const {
    check(flk_ntt_rs_encode_interleaved(dim_, msg.data(), msg.len(), codeword.data(), codeword.len(), num_ntts, stream),
          "rs_encode_interleaved");
  }
  void forward_transform_interleaved(DeviceVec<F128>& data, size_t num_ntts, size_t start_layer = 0,
                                     void* stream = nullptr) const {
    check(flk_ntt_forward_interleaved(dim_, data.data(), data.len(), num_ntts, start_layer, stream),
          "forward_transform_interleaved");
  }

 private:
  explicit AdditiveNttF128(size_t dim) : dim_(dim) {
    if (dim > 24) throw std::invalid_argument("standard NTT: dim <= 24 on the device");
  }
  size_t dim_;
  mutable std::vector<F128> tw_;
};
}  // namespace ntt

// ---------------------------------------------------------------------------
// crate::merkle (CF/merkle.rs:498,683,753) — BLAKE3 trees
// ---------------------------------------------------------------------------
namespace merkle {
enum class HashKind { Sha256 = FLK_HASH_SHA256, Blake3 = FLK_HASH_BLAKE3 };

inline DeviceVec<Hash> merkle_tree(const uint8_t* data, size_t data_len, size_t num_leaves, HashKind kind,
                                   void* stream = nullptr) {
  DeviceVec<Hash> tree(2 * num_leaves - 1, stream);
  check(flk_merkle_tree(data, data_len, num_leaves, static_cast<int>(kind), tree.data()->data(), stream),
        "merkle_tree");
  return tree;
}
}  // namespace merkle

// ---------------------------------------------------------------------------
// crate::pcs (CF/pcs/commit.rs:34-227) and ligerito::ligero_commit (:4750)
// ---------------------------------------------------------------------------
namespace pcs {
constexpr size_t LOG_PACKING = 7;

struct PcsParams {
  size_t m;
  size_t log_inv_rate = 1;
  size_t log_batch_size = 6;
  merkle::HashKind merkle_hash = merkle::HashKind::Blake3;
  size_t log_msg_len() const { return m - LOG_PACKING; }
  size_t log_dim() const { return log_msg_len() - log_batch_size; }
  size_t k_code() const { return log_dim() + log_inv_rate; }
  size_t n_positions() const { return size_t(1) << k_code(); }
  size_t num_ntts() const { return size_t(1) << log_batch_size; }
  size_t codeword_len_f128() const { return n_positions() * num_ntts(); }
  size_t n_leaves() const { return n_positions(); }
};

struct Commitment {
  Hash root;
  PcsParams params;
};

struct ProverData {
  DeviceVec<F128> codeword;
  DeviceVec<Hash> merkle_tree;
};

inline ProverData ligero_commit(const DeviceVec<F128>& poly, size_t log_msg_cols, size_t log_num_interleaved,
                                size_t log_inv_rate, merkle::HashKind kind, Hash* root_out = nullptr,
                                void* stream = nullptr) {
  if (poly.len() != size_t(1) << (log_msg_cols + log_num_interleaved))
    throw std::invalid_argument("ligero_commit: poly length");
  const size_t n_pos = size_t(1) << (log_msg_cols + log_inv_rate);
  ProverData pd{DeviceVec<F128>(n_pos << log_num_interleaved, stream), DeviceVec<Hash>(2 * n_pos - 1, stream)};
  check(flk_commit(poly.data(), log_msg_cols, log_num_interleaved, log_inv_rate, static_cast<int>(kind),
                   pd.codeword.data(), pd.merkle_tree.data()->data(), root_out ? root_out->data() : nullptr, stream),
        "commit");
  return pd;
}

inline std::pair<Commitment, ProverData> commit(const DeviceVec<F128>& z_packed, const PcsParams& params,
                                                void* stream = nullptr) {
  if (z_packed.len() != size_t(1) << params.log_msg_len()) throw std::invalid_argument("commit: z_packed length");
  if (params.log_inv_rate < 1) throw std::invalid_argument("log_inv_rate must be >= 1 for a non-trivial RS code");
  Commitment c{{}, params};
  ProverData pd = ligero_commit(z_packed, params.log_dim(), params.log_batch_size, params.log_inv_rate,
                                params.merkle_hash, &c.root, stream);
  return {c, std::move(pd)};
}
}  // namespace pcs

// ---------------------------------------------------------------------------
// ligerito::SumcheckProver (CF/pcs/ligerito.rs:8893-9280)
// ---------------------------------------------------------------------------
namespace ligerito {
struct SumcheckMessage {
  F128 u_0, u_2;
};

class SumcheckProver {
 public:
  static std::pair<SumcheckProver, SumcheckMessage> new_(DeviceVec<F128> f, DeviceVec<F128> b1, F128 h1,
                                                         void* stream = nullptr) {
    if (f.len() != b1.len()) throw std::invalid_argument("SumcheckProver::new: lengths differ");
    SumcheckProver p(std::move(f), std::move(b1), h1, stream);
    SumcheckMessage msg = p.message_of(p.f_, p.basis_);
    p.transcript_.push_back(msg);
    return {std::move(p), msg};
  }

  SumcheckMessage fold(F128 r) {
    if (pending_) throw std::logic_error("fold before ordinary glue");
    const size_t n = f_.len();
    DeviceVec<F128> nf(n / 2, stream_), nb(n / 2, stream_);
    const std::vector<F128> h = read_reduction(msg_, 2, stream_, [&] {
      check(flk_sc_fold_and_msg(f_.data(), basis_.data(), n, r, nf.data(), nb.data(), msg_.data(), stream_),
            "fold_and_msg");
    });
    f_ = std::move(nf);
    basis_ = std::move(nb);
    SumcheckMessage m{h[0], h[1]};
    transcript_.push_back(m);
    return m;
  }

  SumcheckMessage introduce_new(DeviceVec<F128> b_new, F128 h_new) {
    if (b_new.len() != f_.len()) throw std::invalid_argument("introduce_new: length");
    SumcheckMessage m = message_of(f_, b_new);
    transcript_.push_back(m);
    b_new_ = std::move(b_new);
    h_new_ = h_new;
    pending_ = true;
    return m;
  }

  void glue(F128 alpha) {
    if (!pending_) throw std::logic_error("glue without introduce_new");
    check(flk_sc_glue(basis_.data(), b_new_.data(), basis_.len(), alpha, stream_), "glue");
    t_r_ = t_r_ + alpha * h_new_;
    pending_ = false;
  }

  // The OOD samples of a recursive level (CF/pcs/ligerito.rs:10569-10590) on the
  // transcript `ch` (challenger::FsChallenger): introduce eq(z) with its evaluation,
  // observe (y, u_0, u_2), glue with a sampled beta.  Returns the OOD values.
  template <class Challenger>
  std::vector<F128> ood(Challenger& ch, size_t n_samples) {
    if (pending_) throw std::logic_error("ood before ordinary glue");
    std::vector<F128> ys(n_samples), ms(2 * n_samples);
    check(flk_ligerito_ood(ch.raw(), f_.data(), basis_.data(), log2_exact(f_.len()), n_samples, &t_r_, ys.data(),
                           ms.data(), stream_),
          "ligerito_ood");
    for (size_t i = 0; i < n_samples; ++i) transcript_.push_back(SumcheckMessage{ms[2 * i], ms[2 * i + 1]});
    return ys;
  }

  const DeviceVec<F128>& f() const { return f_; }
  const DeviceVec<F128>& combined_basis() const { return basis_; }
  F128 target() const { return t_r_; }
  const std::vector<SumcheckMessage>& transcript() const { return transcript_; }

 private:
  SumcheckProver(DeviceVec<F128> f, DeviceVec<F128> b, F128 h, void* stream)
      : f_(std::move(f)), basis_(std::move(b)), msg_(2, stream), t_r_(h), stream_(stream) {}
  SumcheckMessage message_of(const DeviceVec<F128>& f, const DeviceVec<F128>& b) {
    const std::vector<F128> h = read_reduction(
        msg_, 2, stream_, [&] { check(flk_sc_round_msg(f.data(), b.data(), f.len(), msg_.data(), stream_), "round_msg"); });
    return SumcheckMessage{h[0], h[1]};
  }
  DeviceVec<F128> f_, basis_, msg_, b_new_;
  F128 t_r_, h_new_{0, 0};
  bool pending_ = false;
  void* stream_;
  std::vector<SumcheckMessage> transcript_;
};
}  // namespace ligerito

// ---------------------------------------------------------------------------
// lincheck / ring_switch / zerocheck reductions
// ---------------------------------------------------------------------------
namespace lincheck {
inline DeviceVec<F128> partial_fold_packed_z(const DeviceVec<uint8_t>& z_packed, size_t m, size_t k_log,
                                             const DeviceVec<F128>& eq_outer, void* stream = nullptr) {
  DeviceVec<F128> out(size_t(1) << k_log, stream);
  check(flk_lc_partial_fold(z_packed.data(), m, k_log, eq_outer.data(), out.data(), stream), "partial_fold");
  return out;
}
}  // namespace lincheck

namespace ring_switch {
inline DeviceVec<F128> fold_1b_rows(const DeviceVec<F128>& packed_witness, const DeviceVec<F128>& suffix_tensor,
                                    void* stream = nullptr) {
  DeviceVec<F128> out(128, stream);
  check(flk_rs_fold_1b_rows(packed_witness.data(), suffix_tensor.data(), packed_witness.len(), out.data(), stream),
        "fold_1b_rows");
  return out;
}
inline DeviceVec<F128> fold_b128_elems(const DeviceVec<F128>& suffix_tensor, const DeviceVec<F128>& eq_r_dprime,
                                       void* stream = nullptr) {
  if (eq_r_dprime.len() != 128) throw std::invalid_argument("fold_b128_elems: eq_r_dprime must have 128 entries");
  DeviceVec<F128> out(suffix_tensor.len(), stream);
  check(flk_rs_fold_b128_elems(suffix_tensor.data(), eq_r_dprime.data(), suffix_tensor.len(), out.data(), stream),
        "fold_b128_elems");
  return out;
}
}  // namespace ring_switch

namespace zerocheck {
inline std::vector<F128> lagrange_weights(size_t k_skip, F128 z) {
  std::vector<F128> w(size_t(1) << k_skip);
  check(flk_zc_lagrange_weights(k_skip, z, w.data()), "lagrange_weights");
  return w;
}
inline std::pair<std::vector<F128>, std::vector<F128>> round1(const DeviceVec<uint8_t>& a, const DeviceVec<uint8_t>& b,
                                                              const DeviceVec<uint8_t>& c, size_t m,
                                                              const std::vector<F128>& r_outer,
                                                              void* stream = nullptr) {
  DeviceVec<F128> out(128, stream);
  check(flk_zc_round1(a.data(), b.data(), c.data(), m, r_outer.data(), out.data(), out.data() + 64, stream), "round1");
  std::vector<F128> h = out.to_host();
  return {std::vector<F128>(h.begin(), h.begin() + 64), std::vector<F128>(h.begin() + 64, h.end())};
}
}  // namespace zerocheck

// ---------------------------------------------------------------------------
// Proof of work and query openings (CF/challenger.rs:384-474,1648;
// CF/merkle.rs:1020-1107)
// ---------------------------------------------------------------------------
namespace challenger {
// pow_scan: the smallest nonce in [start, start + len) or no value.
inline bool pow_scan(const std::array<uint8_t, 32>& digest, uint64_t start, uint64_t len, uint32_t bits,
                     uint64_t* nonce, void* stream = nullptr) {
  uint64_t n = 0;
  check(flk_pow_scan(digest.data(), start, len, bits, &n, stream), "pow_scan");
  *nonce = n;
  return n != ~0ull;
}
}  // namespace challenger

namespace merkle {
// merkle_multi_proof of a device tree (2 num_leaves - 1 hashes), returned to
// the host (the hashes go into the proof).
inline std::vector<Hash> merkle_multi_proof(const DeviceVec<Hash>& tree, size_t num_leaves,
                                            const std::vector<uint64_t>& positions, void* stream = nullptr) {
  size_t depth = 0;
  while ((size_t(1) << depth) < num_leaves) ++depth;
  const size_t cap = positions.size() * (depth ? depth : 1);
  DeviceVec<Hash> tmp(cap ? cap : 1, stream);
  size_t k = 0;
  check(flk_merkle_multi_proof(reinterpret_cast<const uint8_t*>(tree.data()), num_leaves, positions.data(),
                               positions.size(), reinterpret_cast<uint8_t*>(tmp.data()), cap, &k, stream),
        "merkle_multi_proof");
  std::vector<Hash> out(k);
  if (k) {
    check(flk_read_to_host(out.data(), tmp.data(), k * sizeof(Hash), stream), "flk_read_to_host");
  }
  return out;
}
}  // namespace merkle

namespace pcs {
// The opened rows of `positions` from a SoA codeword of rows of row_len elements.
inline DeviceVec<F128> gather_rows(const DeviceVec<F128>& codeword, size_t row_len,
                                   const std::vector<uint64_t>& positions, void* stream = nullptr) {
  if (row_len == 0 || codeword.len() % row_len) throw std::invalid_argument("gather_rows: row_len");
  const size_t n_out = positions.size() * row_len;
  DeviceVec<F128> out(n_out ? n_out : 1, stream);
  check(flk_gather_rows(codeword.data(), codeword.len() / row_len, row_len, positions.data(), positions.size(),
                        out.data(), stream),
        "gather_rows");
  return out;
}
}  // namespace pcs

// ---------------------------------------------------------------------------
// crate::challenger::FsChallenger over BLAKE3 (CF/challenger.rs:226-523):
// host transcript, device nonce search in grind_pow
// ---------------------------------------------------------------------------
namespace challenger {
class FsChallenger {
 public:
  explicit FsChallenger(const std::string& domain) {
    check(flk_fs_new(reinterpret_cast<const uint8_t*>(domain.data()), domain.size(), &t_), "fs_new");
  }
  static FsChallenger with_hash(const std::string& domain) { return FsChallenger(domain); }
  ~FsChallenger() {
    if (t_) flk_fs_free(t_);
  }
  FsChallenger(FsChallenger&& o) noexcept : t_(o.t_) { o.t_ = nullptr; }
  FsChallenger(const FsChallenger& o) { check(flk_fs_clone(o.t_, &t_), "fs_clone"); }
  FsChallenger& operator=(const FsChallenger&) = delete;

  void observe_label(const std::string& l) {
    check(flk_fs_observe_label(t_, reinterpret_cast<const uint8_t*>(l.data()), l.size()), "observe_label");
  }
  void observe_bytes(const uint8_t* b, size_t n) { check(flk_fs_observe_bytes(t_, b, n), "observe_bytes"); }
  void observe_f128(F128 v) { check(flk_fs_observe_f128(t_, v), "observe_f128"); }
  void observe_f128_slice(const std::vector<F128>& v) {
    check(flk_fs_observe_f128_slice(t_, v.data(), v.size()), "observe_f128_slice");
  }
  F128 sample_f128() {
    F128 v;
    check(flk_fs_sample_f128(t_, &v), "sample_f128");
    return v;
  }
  std::vector<F128> sample_f128_vec(size_t n) {
    std::vector<F128> v(n);
    check(flk_fs_sample_f128_vec(t_, n, v.data()), "sample_f128_vec");
    return v;
  }
  std::array<uint8_t, 32> state_digest() const {
    std::array<uint8_t, 32> d;
    check(flk_fs_state_digest(t_, d.data()), "state_digest");
    return d;
  }
  uint64_t grind_pow(uint32_t bits, void* stream = nullptr) {
    uint64_t n = 0;
    check(flk_fs_grind_pow(t_, bits, &n, stream), "grind_pow");
    return n;
  }
  bool verify_pow(uint64_t nonce, uint32_t bits) {
    int ok = 0;
    check(flk_fs_verify_pow(t_, nonce, bits, &ok), "verify_pow");
    return ok != 0;
  }
  std::vector<uint64_t> sample_distinct_queries(uint64_t block_len, size_t count) {
    std::vector<uint64_t> q(count);
    check(flk_fs_sample_distinct_queries(t_, block_len, count, q.data()), "sample_distinct_queries");
    return q;
  }
  flk_fs* raw() { return t_; }

 private:
  flk_fs* t_ = nullptr;
};
}  // namespace challenger

// ---------------------------------------------------------------------------
// ring_switch::prove_batched (CF/pcs/ring_switch.rs:3727-3735, 4310-4383) and
// pcs::compute_combined_basis_and_target (CF/pcs.rs:778-870) for ring-switched
// claims; the combined basis (the sum of the gamma-baked rs_eq_ind) stays on the
// device.  x_outers: one vector of log2(len) + 1 coordinates per claim.
// ---------------------------------------------------------------------------
struct RingSwitchBatch {
  std::vector<std::vector<F128>> s_hat_v;
  std::vector<F128> sumcheck_claims, gammas;
  DeviceVec<F128> b_combined;
  F128 target_combined{0, 0};
};

namespace detail {
inline RingSwitchBatch ring_switch_call(const DeviceVec<F128>& z, const std::vector<std::vector<F128>>& x_outers,
                                        challenger::FsChallenger& ch, bool want_basis, bool combine,
                                        void* stream) {
  const size_t log_len = log2_exact(z.len()), k = x_outers.size();
  std::vector<F128> xs;
  for (const auto& x : x_outers) {
    if (x.size() != log_len + 1) throw std::invalid_argument("ring switch: x_outer needs log2(len) + 1 coordinates");
    xs.insert(xs.end(), x.begin(), x.end());
  }
  RingSwitchBatch out;
  std::vector<F128> shv(128 * k);
  out.sumcheck_claims.resize(k);
  out.gammas.resize(k);
  if (want_basis || combine) out.b_combined = DeviceVec<F128>(z.len(), stream);
  if (combine)
    check(flk_pcs_combined_basis_and_target(ch.raw(), z.data(), log_len, xs.data(), k, shv.data(),
                                            out.sumcheck_claims.data(), out.gammas.data(), out.b_combined.data(),
                                            &out.target_combined, stream),
          "pcs_combined_basis_and_target");
  else
    check(flk_rs_prove_batched(ch.raw(), z.data(), log_len, xs.data(), k, shv.data(), out.sumcheck_claims.data(),
                               out.gammas.data(), want_basis ? out.b_combined.data() : nullptr, stream),
          "rs_prove_batched");
  for (size_t j = 0; j < k; ++j) out.s_hat_v.emplace_back(shv.begin() + 128 * j, shv.begin() + 128 * (j + 1));
  return out;
}
}  // namespace detail

namespace ring_switch {
inline RingSwitchBatch prove_batched(const DeviceVec<F128>& packed_witness,
                                     const std::vector<std::vector<F128>>& x_outers, challenger::FsChallenger& ch,
                                     bool want_basis = true, void* stream = nullptr) {
  return detail::ring_switch_call(packed_witness, x_outers, ch, want_basis, false, stream);
}
}  // namespace ring_switch

namespace pcs {
inline RingSwitchBatch compute_combined_basis_and_target(const DeviceVec<F128>& packed_witness,
                                                         const std::vector<std::vector<F128>>& x_outers,
                                                         challenger::FsChallenger& ch, void* stream = nullptr) {
  return detail::ring_switch_call(packed_witness, x_outers, ch, true, true, stream);
}
}  // namespace pcs

// ---------------------------------------------------------------------------
// flock_prover::r1cs_hashes::blake3 witness generation
// (crates/flock-prover/src/r1cs_hashes/blake3.rs:2184-2206,4617-4647)
// ---------------------------------------------------------------------------
namespace r1cs_hashes {
namespace blake3 {
using Compression = flk_compression;  // (cv, m, counter, block_len, flags)

struct Witness {
  DeviceVec<F128> z, a, b;       // 2^n_blocks_log x 128 elements each
  DeviceVec<uint8_t> z_lincheck;  // 2^n_blocks_log x 2048 bytes (empty without the stripe)
};

inline Witness generate_witness_with_ab_packed_and_lincheck(const DeviceVec<Compression>& blocks,
                                                            size_t n_blocks_log, bool lincheck = true,
                                                            void* stream = nullptr) {
  if (n_blocks_log < 3 || n_blocks_log > 30)
    throw std::invalid_argument("generate_witness: n_blocks_log must be in 3..30");
  const size_t n_total = size_t(1) << n_blocks_log;
  if (blocks.len() > n_total) throw std::invalid_argument("generate_witness: more compressions than slots");
  Witness w{DeviceVec<F128>(n_total * 128, stream), DeviceVec<F128>(n_total * 128, stream),
            DeviceVec<F128>(n_total * 128, stream),
            lincheck ? DeviceVec<uint8_t>(n_total * 2048, stream) : DeviceVec<uint8_t>()};
  check(flk_blake3_witness(blocks.data(), blocks.len(), n_blocks_log, w.z.data(), w.a.data(), w.b.data(),
                           lincheck ? w.z_lincheck.data() : nullptr, stream),
        "generate_witness");
  return w;
}
inline Witness generate_witness_with_ab_packed(const DeviceVec<Compression>& blocks, size_t n_blocks_log,
                                               void* stream = nullptr) {
  return generate_witness_with_ab_packed_and_lincheck(blocks, n_blocks_log, false, stream);
}
}  // namespace blake3
}  // namespace r1cs_hashes

// ---------------------------------------------------------------------------
// recursive_prover_with_basis (CF/pcs/ligerito.rs:9996-10026, body :10185-11040)
// sequenced over the device entry points, in the transcript order its verifier
// replays (:11836-12228).  Grinds return the smallest valid nonce.
// ---------------------------------------------------------------------------
namespace ligerito {
struct ProverConfig {
  size_t initial_k;
  std::vector<size_t> recursive_ks, log_inv_rates, queries, ood_samples;
  std::vector<uint32_t> grinding_bits, fold_grinding_bits;
};

struct OpenedLevel {
  std::vector<F128> rows;  // n_queries x row_len, query-major
  std::vector<Hash> merkle_proof;
};

struct LigeritoProof {
  Hash initial_root;
  OpenedLevel initial_proof;
  std::vector<Hash> recursive_roots;
  std::vector<OpenedLevel> recursive_proofs;
  std::vector<F128> yr;
  OpenedLevel final_proof;
  std::vector<SumcheckMessage> sumcheck_transcript;
  std::vector<uint64_t> grinding_nonces, fold_grinding_nonces;
  std::vector<F128> ood_values;
};

inline size_t ceil_log2(size_t n) {
  size_t k = 0;
  while ((size_t(1) << k) < n) ++k;
  return k;
}

inline LigeritoProof recursive_prover_with_basis(const ProverConfig& cfg, DeviceVec<F128> f, DeviceVec<F128> b,
                                                 F128 target, const DeviceVec<F128>& l0_codeword,
                                                 const DeviceVec<Hash>& l0_tree, challenger::FsChallenger& ch,
                                                 void* stream = nullptr) {
  const size_t r = cfg.recursive_ks.size(), k0 = cfg.initial_k;
  if (r < 1 || cfg.log_inv_rates.size() != r + 1 || cfg.queries.size() != r + 1 || cfg.ood_samples.size() != r + 1 ||
      cfg.grinding_bits.size() != r + 1 || cfg.fold_grinding_bits.size() != r + 1 || cfg.ood_samples[0])
    throw std::invalid_argument("recursive_prover_with_basis: config");
  if (f.len() != b.len()) throw std::invalid_argument("recursive_prover_with_basis: f and b lengths");
  const size_t log_n = log2_exact(f.len());
  size_t folds = k0;
  for (size_t k : cfg.recursive_ks) folds += k;
  // Every fold needs two elements per output and the final yr at least one variable.
  if (folds + 1 > log_n) throw std::invalid_argument("recursive_prover_with_basis: the config folds all variables");
  size_t n_cur = log_n - k0;
  const size_t block0 = size_t(1) << (n_cur + cfg.log_inv_rates[0]);
  if (l0_codeword.len() != block0 << k0) throw std::invalid_argument("recursive_prover_with_basis: L0 codeword");
  LigeritoProof proof;
  proof.initial_root = l0_tree.back();
  auto observe_msg = [&](const SumcheckMessage& m) {
    ch.observe_f128(m.u_0);
    ch.observe_f128(m.u_2);
  };
  // One opening of a level and, for every level but the last, the basis it
  // induces: the rows are gathered on the device and used there by the
  // induction, the multi-proof hashes gathered beside them, and rows, hashes
  // and the induced sum come back in one read-back (the index lists and
  // constants go up through the staged uploads, which do not wait).
  struct Opening {
    OpenedLevel level;
    DeviceVec<F128> basis;
    F128 h{0, 0};
  };
  auto open = [&](const DeviceVec<F128>& cw, const DeviceVec<Hash>& tree, size_t row_len,
                  const std::vector<uint64_t>& q, size_t log_cols, const std::vector<F128>* v,
                  const std::vector<F128>* alpha) {
    const size_t n_rows = cw.len() / row_len, nr = q.size() * row_len;
    size_t depth = 0;
    while ((size_t(1) << depth) < n_rows) ++depth;
    const size_t cap = q.size() * (depth ? depth : 1);
    // [rows (nr F128) | induced sum (1 F128) | hashes (cap x 32 B)] in one device buffer
    const size_t n128 = nr + 1 + 2 * cap;
    DeviceVec<F128> buf(n128, stream);
    F128* rows = buf.data();
    F128* sum = rows + nr;
    uint8_t* hashes = reinterpret_cast<uint8_t*>(sum + 1);
    check(flk_gather_rows(cw.data(), n_rows, row_len, q.data(), q.size(), rows, stream), "gather_rows");
    size_t k = 0;
    check(flk_merkle_multi_proof(reinterpret_cast<const uint8_t*>(tree.data()), n_rows, q.data(), q.size(), hashes,
                                 cap, &k, stream),
          "merkle_multi_proof");
    Opening out;
    if (v) {
      DeviceVec<uint64_t> qd(q, stream);
      out.basis = DeviceVec<F128>(size_t(1) << log_cols, stream);
      check(flk_induce_sumcheck_poly(log_cols, rows, q.size(), v->data(), v->size(), qd.data(), alpha->data(),
                                     alpha->size(), out.basis.data(), sum, stream),
            "induce_sumcheck_poly");
    }
    const size_t n_read = nr + 1 + 2 * k;
    std::vector<F128> host(n_read);
    check(flk_read_to_host(host.data(), buf.data(), n_read * sizeof(F128), stream), "flk_read_to_host");
    out.level.rows.assign(host.begin(), host.begin() + nr);
    out.h = host[nr];
    out.level.merkle_proof.resize(k);
    if (k) memcpy(out.level.merkle_proof.data(), host.data() + nr + 1, k * sizeof(Hash));
    return out;
  };

  ch.observe_label("flock-ligerito-basis-v0");
  ch.observe_f128(target);
  ch.observe_bytes(proof.initial_root.data(), 32);
  auto started = SumcheckProver::new_(std::move(f), std::move(b), target, stream);
  SumcheckProver& sp = started.first;
  observe_msg(started.second);
  auto fold_rounds = [&](size_t k, size_t level) {
    std::vector<F128> rs;
    for (size_t j = 0; j < k; ++j) {
      const uint32_t fb = cfg.fold_grinding_bits[level];
      if (fb > j) proof.fold_grinding_nonces.push_back(ch.grind_pow(fb - (uint32_t)j, stream));
      const F128 ri = ch.sample_f128();
      rs.push_back(ri);
      observe_msg(sp.fold(ri));
    }
    return rs;
  };
  auto introduce = [&](DeviceVec<F128> basis, F128 h) {
    observe_msg(sp.introduce_new(std::move(basis), h));
    sp.glue(ch.sample_f128());
  };
  auto ood = [&](size_t level) {
    auto ys = sp.ood(ch, cfg.ood_samples[level]);
    proof.ood_values.insert(proof.ood_values.end(), ys.begin(), ys.end());
  };

  const std::vector<F128> r_lane = fold_rounds(k0, 0);
  size_t k_next = cfg.recursive_ks[0];
  pcs::ProverData prev = pcs::ligero_commit(sp.f(), n_cur - k_next, k_next, cfg.log_inv_rates[1],
                                            merkle::HashKind::Blake3, nullptr, stream);
  size_t prev_k = k_next, prev_block = size_t(1) << (n_cur - k_next + cfg.log_inv_rates[1]);
  proof.recursive_roots.push_back(prev.merkle_tree.back());
  ch.observe_bytes(proof.recursive_roots.back().data(), 32);
  ood(1);
  proof.grinding_nonces.push_back(ch.grind_pow(cfg.grinding_bits[0], stream));
  auto q0 = ch.sample_distinct_queries(block0, cfg.queries[0]);
  auto a0 = ch.sample_f128_vec(ceil_log2(cfg.queries[0]));
  {
    Opening o = open(l0_codeword, l0_tree, size_t(1) << k0, q0, n_cur, &r_lane, &a0);
    proof.initial_proof = std::move(o.level);
    introduce(std::move(o.basis), o.h);
  }
  for (size_t i = 0; i < r; ++i) {
    const std::vector<F128> level_rs = fold_rounds(cfg.recursive_ks[i], i + 1);
    n_cur -= cfg.recursive_ks[i];
    if (i == r - 1) {
      proof.yr = sp.f().to_host();
      for (const F128& v : proof.yr) ch.observe_f128(v);
      proof.grinding_nonces.push_back(ch.grind_pow(cfg.grinding_bits[i + 1], stream));
      auto ql = ch.sample_distinct_queries(prev_block, cfg.queries[i + 1]);
      proof.final_proof = open(prev.codeword, prev.merkle_tree, size_t(1) << prev_k, ql, 0, nullptr, nullptr).level;
      proof.sumcheck_transcript = sp.transcript();
      return proof;
    }
    k_next = cfg.recursive_ks[i + 1];
    pcs::ProverData nxt = pcs::ligero_commit(sp.f(), n_cur - k_next, k_next, cfg.log_inv_rates[i + 2],
                                             merkle::HashKind::Blake3, nullptr, stream);
    proof.recursive_roots.push_back(nxt.merkle_tree.back());
    ch.observe_bytes(proof.recursive_roots.back().data(), 32);
    ood(i + 2);
    proof.grinding_nonces.push_back(ch.grind_pow(cfg.grinding_bits[i + 1], stream));
    auto qi = ch.sample_distinct_queries(prev_block, cfg.queries[i + 1]);
    auto ai = ch.sample_f128_vec(ceil_log2(cfg.queries[i + 1]));
    {
      Opening o = open(prev.codeword, prev.merkle_tree, size_t(1) << prev_k, qi, n_cur, &level_rs, &ai);
      proof.recursive_proofs.push_back(std::move(o.level));
      introduce(std::move(o.basis), o.h);
    }
    prev = std::move(nxt);
    prev_k = k_next;
    prev_block = size_t(1) << (n_cur - k_next + cfg.log_inv_rates[i + 2]);
  }
  throw std::logic_error("recursive_prover_with_basis: unreachable");
}
}  // namespace ligerito

}  // namespace flock
