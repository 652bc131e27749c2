// The BLAKE3 R1CS prover sequenced over the device entry points
// (CF = crates/flock-core/src/, CP = crates/flock-prover/src/):
//
//   flk_zerocheck_prove   zerocheck::prove_packed_padded (CF/zerocheck.rs:384-535,
//                         body :539-1620; transcript order of verify :1642-1790)
//   flk_lincheck_prove    lincheck::prove_padded (CF/lincheck.rs:5345-5878)
//   flk_blake3_setup_*    Blake3Setup::new (CP/r1cs_hashes/blake3.rs:4706-4788) with
//                         the Ligerito ladder of (m, profile) (CF/pcs/ligerito.rs:256-345)
//   flk_blake3_prove      witness + pcs::commit + prover::prove_ligerito
//                         (CP/prover.rs:373-472) + R1csProofBundleLigerito::to_bytes
//                         (CP/proof_io.rs:160-198)
//
// Every data-parallel step runs on the device on the caller's stream; the
// host keeps the Fiat-Shamir transcript and the O(1)-sized protocol algebra
// (Lagrange weights over 64 nodes, the 64-point interpolation of round 1's C
// message, the final lincheck combination).  Each round reads its message
// back (a few F128) before the next challenge can be sampled; that read-back
// is the only synchronisation.
#include <algorithm>
#include <array>
#include <chrono>
#include <cstring>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/flock_gpu.h"
#include "../../include/flock_host.hpp"
#include "host_api.hpp"
#include "r1cs.hpp"

using flock::DeviceVec;
using flock::F128;
using flock::Hash;

namespace {

constexpr size_t K_SKIP = flk::r1cs::K_SKIP;
constexpr size_t K_LOG = flk::r1cs::K_LOG;
constexpr size_t N_INNER = 7;  // the fixed small + medium eq coordinates of round 1

void chk(int rc, const char* where) {
  if (rc != FLK_OK) throw flock::GpuError(rc, where);
}

struct Fs {
  flk_fs* t;
  void label(const char* l) { chk(flk_fs_observe_label(t, reinterpret_cast<const uint8_t*>(l), std::strlen(l)), "observe_label"); }
  void f128(F128 v) { chk(flk_fs_observe_f128(t, v), "observe_f128"); }
  void slice(const std::vector<F128>& v) { chk(flk_fs_observe_f128_slice(t, v.data(), v.size()), "observe_f128_slice"); }
  void bytes(const uint8_t* b, size_t n) { chk(flk_fs_observe_bytes(t, b, n), "observe_bytes"); }
  F128 sample() {
    F128 v;
    chk(flk_fs_sample_f128(t, &v), "sample_f128");
    return v;
  }
  std::vector<F128> sample_vec(size_t n) {
    std::vector<F128> v(n);
    chk(flk_fs_sample_f128_vec(t, n, v.data()), "sample_f128_vec");
    return v;
  }
};

F128 inner(const std::vector<F128>& a, const std::vector<F128>& b) {
  F128 acc{0, 0};
  for (size_t i = 0; i < a.size(); ++i) acc = flk::fadd(acc, flk::host_mul(a[i], b[i]));
  return acc;
}

// ---------------------------------------------------------------------------
// Zerocheck
// ---------------------------------------------------------------------------
struct ZcProof {
  std::vector<F128> round1_ab, round1_c;
  std::vector<std::array<F128, 2>> rounds;  // (G(1), G(inf)) per multilinear round
  F128 final_a, final_b, final_c;
};
struct ZcClaim {
  F128 z;
  std::vector<F128> mlv, r_rest;  // m - 6 each
};

// s_c (device, 128 entries, or NULL) receives round 1's column sums of c, the
// ring-switch s_hat_v of the c claim (its suffix point is r[7..m), the same eq).
void zerocheck_prove(Fs& ch, const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                     ZcProof& pf, ZcClaim& cl, void* st, F128* s_c = nullptr) {
  const size_t n_mlv = m - K_SKIP;
  ch.label("flock-zerocheck-v0");
  // r = [r_skip (6), small (3), medium (4), r_outer (m - 13)] (CF/zerocheck.rs:576-586).
  std::vector<F128> r = ch.sample_vec(K_SKIP);
  std::vector<F128> r_outer = ch.sample_vec(m - K_SKIP - N_INNER);
  F128 fixed[N_INNER];
  flk::zc_fixed_challenges(fixed);
  for (const F128& f : fixed) r.push_back(f);
  r.insert(r.end(), r_outer.begin(), r_outer.end());

  // Round 1 in the naive convention (the C_s factor is part of round1_naive).
  DeviceVec<F128> r1(128, st);
  chk(flk::zc_round1_keep_s(a_bits, b_bits, c_bits, m, r_outer.data(), r1.data(), r1.data() + 64, s_c, st),
      "zc_round1");
  // The eq tables of the device multilinear rounds depend on r alone, so they are
  // enqueued behind round 1 instead of after its read-back.  eq_lo: each device
  // round's first 8 eq coordinates; eq_hi (two buffers, alternating): the rest,
  // handed from round to round by summing out the lowest coordinate.
  const size_t d0 = n_mlv - 1;
  const size_t n_dev = d0 >= 8 ? d0 - 7 : 0;  // rounds i with n_mlv - i - 1 >= 8
  DeviceVec<F128> eq_lo(256 * (n_dev ? n_dev : 1), st),
      eq_hi[2] = {DeviceVec<F128>(d0 >= 8 ? size_t(1) << (d0 - 8) : 1, st),
                  DeviceVec<F128>(d0 >= 9 ? size_t(1) << (d0 - 9) : 1, st)};
  for (size_t i = 0; i < n_dev; ++i)
    chk(flk_build_eq(r.data() + K_SKIP + i + 1, 8, eq_lo.data() + 256 * i, st), "build_eq(lo)");
  if (d0 >= 8) chk(flk_build_eq(r.data() + K_SKIP + 1 + 8, d0 - 8, eq_hi[0].data(), st), "build_eq(hi)");
  const std::vector<F128> h = r1.to_host();
  pf.round1_ab.assign(h.begin(), h.begin() + 64);
  pf.round1_c.assign(h.begin() + 64, h.end());
  ch.slice(pf.round1_ab);
  ch.slice(pf.round1_c);
  const F128 z = ch.sample();

  // c_eval = P^C(z) by interpolation on Lambda (interpolate_at_z_on_lambda, multilinear.rs:410-428).
  std::vector<F128> lam(64), ws(64);
  chk(flk::lagrange_weights_nodes(K_SKIP, z, 64, lam.data()), "lagrange_weights(Lambda)");
  pf.final_c = inner(lam, pf.round1_c);

  // Fold at z on S, then the multilinear rounds binding the low variable first
  // (fold_at_z_naive, round_pair_naive with r_next = [1, r[k_skip + i + 1 ..]],
  // fold_in_place_pair; multilinear.rs:575-640, 3194-3213).  While the round's eq
  // has >= 8 coordinates the device runs each round as one fused pass: the bind by
  // the previous challenge, the new tables written, and the message against the eq
  // split into its first 8 coordinates and the rest (flk_zc_fold_round).  Rounds 0
  // and 1 read the a and b bits themselves (flk_zc_fold_at_z_round without tables,
  // flk_zc_fold_at_z_bind_round), so the 2^(m-6)-entry fold at z is never stored.
  // The last rounds (tables of <= 512 entries) run on the host.
  chk(flk::lagrange_weights_nodes(K_SKIP, z, 0, ws.data()), "lagrange_weights(S)");
  DeviceVec<F128> wd(ws, st);
  size_t n = size_t(1) << n_mlv;
  DeviceVec<F128> am, bm, msg(2, st);
  // With two device rounds, rounds 0 and 1 both read the bits and the fold at z is
  // never written (round 0 sends the message alone, round 1 writes the table
  // bound by rho_0); otherwise the fold at z is written in full.
  const bool from_bits = n_dev >= 2;
  if (!from_bits) {
    am = DeviceVec<F128>(n, st);
    bm = DeviceVec<F128>(n, st);
  }
  if (d0 < 8) {
    chk(flk_zc_fold_at_z(a_bits, m, K_SKIP, wd.data(), am.data(), st), "zc_fold_at_z(a)");
    chk(flk_zc_fold_at_z(b_bits, m, K_SKIP, wd.data(), bm.data(), st), "zc_fold_at_z(b)");
  }
  std::vector<F128> rhos, ha, hb;  // ha, hb: the host tail's tables
  bool on_device = true;
  auto host_fold = [&](std::vector<F128>& t, F128 rho) {
    std::vector<F128> o(t.size() / 2);
    for (size_t j = 0; j < o.size(); ++j)
      o[j] = flk::fadd(t[2 * j], flk::host_mul(rho, flk::fadd(t[2 * j], t[2 * j + 1])));
    t.swap(o);
  };
  for (size_t i = 0; i < n_mlv; ++i) {
    const size_t d = n_mlv - i - 1;  // eq over r[k_skip + i + 1 .. m)
    const F128* pt = r.data() + K_SKIP + i + 1;
    std::array<F128, 2> g;
    if (on_device && d >= 8) {
      DeviceVec<F128>& hi = eq_hi[i & 1];
      DeviceVec<F128>& hi_next = eq_hi[(i + 1) & 1];
      const F128* lo = eq_lo.data() + 256 * i;  // eq over pt[0 .. 8)
      // each round's message comes back through an armed read-back: the round
      // kernel's last work-group stores it into host memory
      std::vector<F128> h;
      if (i == 0) {  // the fold at z and the first message in one pass over the bits
        h = flock::read_reduction(msg, 2, st, [&] {
          chk(flk_zc_fold_at_z_round(a_bits, b_bits, m, wd.data(), lo, hi.data(), from_bits ? nullptr : am.data(),
                                     from_bits ? nullptr : bm.data(), hi_next.data(), msg.data(), st),
              "zc_fold_at_z_round");
        });
      } else if (i == 1 && from_bits) {  // the fold at z bound by rho_0, and the second message
        am = DeviceVec<F128>(n / 2, st);
        bm = DeviceVec<F128>(n / 2, st);
        h = flock::read_reduction(msg, 2, st, [&] {
          chk(flk_zc_fold_at_z_bind_round(a_bits, b_bits, m, wd.data(), &rhos.back(), lo, hi.data(), am.data(),
                                          bm.data(), hi_next.data(), msg.data(), st),
              "zc_fold_at_z_bind_round");
        });
        n /= 2;
      } else {
        DeviceVec<F128> an(n / 2, st), bn(n / 2, st);
        h = flock::read_reduction(msg, 2, st, [&] {
          chk(flk_zc_fold_round(am.data(), bm.data(), n, &rhos.back(), lo, hi.data(), d - 8, an.data(), bn.data(),
                                hi_next.data(), msg.data(), st),
              "zc_fold_round");
        });
        am = std::move(an);
        bm = std::move(bn);
        n /= 2;
      }
      g = {h[0], h[1]};
    } else {
      if (on_device) {  // the tables move to the host once they are small
        ha = am.to_host();
        hb = bm.to_host();
        am = DeviceVec<F128>();
        bm = DeviceVec<F128>();
        on_device = false;
      }
      if (i) {
        host_fold(ha, rhos.back());
        host_fold(hb, rhos.back());
      }
      std::vector<F128> eq(size_t(1) << d);
      eq[0] = F128{1, 0};
      for (size_t j = 0; j < d; ++j)
        for (size_t k = 0; k < (size_t(1) << j); ++k) {
          eq[k + (size_t(1) << j)] = flk::host_mul(eq[k], pt[j]);
          eq[k] = flk::host_mul(eq[k], F128{pt[j].lo ^ 1, pt[j].hi});
        }
      F128 g1{0, 0}, gi{0, 0};
      for (size_t x = 0; x < eq.size(); ++x) {
        const F128 a0 = ha[2 * x], a1 = ha[2 * x + 1], b0 = hb[2 * x], b1 = hb[2 * x + 1];
        g1 = flk::fadd(g1, flk::host_mul(eq[x], flk::host_mul(a1, b1)));
        gi = flk::fadd(gi, flk::host_mul(eq[x], flk::host_mul(flk::fadd(a0, a1), flk::fadd(b0, b1))));
      }
      g = {g1, gi};
    }
    pf.rounds.push_back(g);
    ch.f128(g[0]);
    ch.f128(g[1]);
    rhos.push_back(ch.sample());
  }
  if (on_device) {  // only when no round ran on the host (m - 6 rounds all with >= 8 eq coordinates: never)
    ha = am.to_host();
    hb = bm.to_host();
  }
  host_fold(ha, rhos.back());
  host_fold(hb, rhos.back());
  pf.final_a = ha[0];
  pf.final_b = hb[0];
  ch.f128(pf.final_a);
  ch.f128(pf.final_b);
  cl.z = z;
  cl.mlv = rhos;
  cl.r_rest.assign(r.begin() + K_SKIP, r.end());
}

// ---------------------------------------------------------------------------
// Lincheck
// ---------------------------------------------------------------------------
struct LcProof {
  std::vector<std::array<F128, 2>> rounds;  // (q(1), q(inf)) per round
  std::vector<F128> z_partial;              // 2^k_skip
};
struct LcClaim {
  F128 r_inner_skip;
  std::vector<F128> r_inner_rest;
  F128 w;
};

// x_ab = z_skip || x_inner_rest (k_log - k_skip) || x_outer (m - k_log).
void lincheck_prove(Fs& ch, const flk_r1cs* r1cs, const uint8_t* z_stripe, size_t m, const F128* x_ab, LcProof& pf,
                    LcClaim& cl, void* st) {
  const size_t k = size_t(1) << K_LOG, rest = K_LOG - K_SKIP, n_outer = m - K_LOG;
  ch.label("flock-lincheck-v0");
  const F128 alpha = ch.sample();
  const F128 beta = ch.sample();  // the constant-wire pin, sampled right after alpha
  DeviceVec<F128> comb(k, st), zv(k, st), msg(2, st);
  chk(flk_lincheck_comb(r1cs, x_ab[0], x_ab + 1, rest, alpha, beta, comb.data(), st), "lincheck_comb");
  {
    DeviceVec<F128> eq_outer(size_t(1) << n_outer, st);
    chk(flk_build_eq(x_ab + 1 + rest, n_outer, eq_outer.data(), st), "build_eq(x_outer)");
    chk(flk_lc_partial_fold(z_stripe, m, K_LOG, eq_outer.data(), zv.data(), st), "lc_partial_fold");
  }
  size_t n = k;
  std::vector<F128> rr;
  for (size_t t = 0; t < rest; ++t) {
    const std::vector<F128> q = flock::read_reduction(
        msg, 2, st, [&] { chk(flk_lc_round_msg_msb(comb.data(), zv.data(), n, msg.data(), st), "lc_round_msg_msb"); });
    ch.f128(q[0]);
    ch.f128(q[1]);
    const F128 ri = ch.sample();
    pf.rounds.push_back({q[0], q[1]});
    rr.push_back(ri);
    chk(flk_lc_bind_top(comb.data(), n, ri, st), "lc_bind_top(comb)");
    chk(flk_lc_bind_top(zv.data(), n, ri, st), "lc_bind_top(z)");
    n /= 2;
  }
  pf.z_partial.resize(n);
  chk(flk_read_to_host(pf.z_partial.data(), zv.data(), n * sizeof(F128), st), "read z_partial");
  ch.slice(pf.z_partial);
  cl.r_inner_skip = ch.sample();
  std::vector<F128> lam(size_t(1) << K_SKIP);
  chk(flk::lagrange_weights_nodes(K_SKIP, cl.r_inner_skip, 0, lam.data()), "lagrange_weights");
  cl.w = inner(lam, pf.z_partial);
  cl.r_inner_rest.assign(rr.rbegin(), rr.rend());  // LSB-first order of x_inner_rest
}

// ---------------------------------------------------------------------------
// Ligerito ladders (generated data; CF/pcs/ligerito.rs:963-1013)
// ---------------------------------------------------------------------------
struct Level {
  uint32_t log_inv_rate, log_msg_cols, log_num_interleaved, k_recursive, queries, grinding_bits,
      fold_grinding_bits, ood_samples;
};
struct Ladder {
  uint32_t m, profile, log_n, initial_k, yr_log_n, n_levels;
  Level levels[8];
};
const Ladder LADDERS[] = {
#include "ligerito_ladders.inc"
};

const Ladder* find_ladder(size_t m, int profile) {
  for (const Ladder& l : LADDERS)
    if (l.m == m && (int)l.profile == profile) return &l;
  return nullptr;
}

flock::ligerito::ProverConfig prover_config(const Ladder& l) {
  flock::ligerito::ProverConfig c;
  c.initial_k = l.initial_k;
  for (uint32_t i = 0; i < l.n_levels; ++i) {
    const Level& v = l.levels[i];
    if (i) c.recursive_ks.push_back(v.k_recursive);
    c.log_inv_rates.push_back(v.log_inv_rate);
    c.queries.push_back(v.queries);
    c.ood_samples.push_back(v.ood_samples);
    c.grinding_bits.push_back(v.grinding_bits);
    c.fold_grinding_bits.push_back(v.fold_grinding_bits);
  }
  return c;
}

// ---------------------------------------------------------------------------
// bincode fixint encoding of R1csProofBundleLigerito (CP/proof_io.rs:160-198,
// 804-941): u64le lengths, F128 = lo || hi, Hash = 32 raw bytes, unit enums
// as u32le variant indices, structs as field concatenation.
// ---------------------------------------------------------------------------
struct Writer {
  std::vector<uint8_t> b;
  void raw(const void* p, size_t n) {
    const uint8_t* q = static_cast<const uint8_t*>(p);
    b.insert(b.end(), q, q + n);
  }
  void u64(uint64_t v) {
    for (int i = 0; i < 8; ++i) b.push_back(uint8_t(v >> (8 * i)));
  }
  void u32(uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(uint8_t(v >> (8 * i)));
  }
  void f128(F128 v) {
    u64(v.lo);
    u64(v.hi);
  }
  void f128_vec(const F128* v, size_t n) {
    u64(n);
    for (size_t i = 0; i < n; ++i) f128(v[i]);
  }
  void f128_vec(const std::vector<F128>& v) { f128_vec(v.data(), v.size()); }
  void pairs(const std::vector<std::array<F128, 2>>& v) {
    u64(v.size());
    for (const auto& p : v) {
      f128(p[0]);
      f128(p[1]);
    }
  }
  void hashes(const std::vector<Hash>& v) {
    u64(v.size());
    for (const Hash& h : v) raw(h.data(), 32);
  }
  void u64_vec(const std::vector<uint64_t>& v) {
    u64(v.size());
    for (uint64_t x : v) u64(x);
  }
  // Vec<Vec<F128>> of n_rows rows of row_len elements (query-major flat rows).
  void rows(const std::vector<F128>& flat, size_t row_len) {
    const size_t n_rows = row_len ? flat.size() / row_len : 0;
    u64(n_rows);
    for (size_t q = 0; q < n_rows; ++q) f128_vec(flat.data() + q * row_len, row_len);
  }
};

}  // namespace

struct flk_blake3_setup {
  size_t n_blocks_log, m;
  int profile;
  const Ladder* ladder;
  flk_r1cs* r1cs;
  uint8_t digest[32];
  // Device arenas for the proves of this setup, one per concurrent prove: created
  // on demand, reused by later proves (no device allocation per prove).
  size_t arena_bytes = 0;
  std::mutex arena_mu;
  std::vector<flk::Arena*> arenas, idle;
};

namespace {

// Bytes of device memory one prove holds at its peak (the zerocheck rounds):
// z, a, b and the stripe (2^(m-3) bytes each), the L0 codeword (2^(m-3+rate))
// and tree (64 * 2^(m-13+rate)), the a, b tables bound by rho_0 and their halves
// (3 * 2^(m-3); round 0 stores no table), plus 1/8 headroom for the smaller
// tables.  At m = 32 (fast) the measured peak is 4.90 GB against 5.78 GB here.
size_t prove_footprint(size_t m, size_t rate) {
  const size_t z = size_t(1) << (m - 3);
  const size_t base = 4 * z + (z << rate) + 64 * (size_t(1) << (m - 13 + rate)) + 3 * z;
  return base + base / 8;
}

int guarded(const char* where, int (*fn)(void*), void* ctx) {
  try {
    return fn(ctx);
  } catch (const flock::GpuError& e) {
    return e.code;
  } catch (const std::invalid_argument& e) {
    return flk::set_error(FLK_E_SHAPE, "%s: %s", where, e.what());
  } catch (const std::bad_alloc&) {
    return flk::set_error(FLK_E_NOMEM, "%s: out of host memory", where);
  } catch (const std::exception& e) {
    return flk::set_error(FLK_E_SHAPE, "%s: %s", where, e.what());
  }
}

struct ProveCall {
  const flk_blake3_setup* s;
  const flk_compression* blocks;
  size_t n_blocks;
  const uint8_t* domain;
  size_t domain_len;
  uint8_t* out;
  size_t cap;
  size_t* out_len;
  double* stage_s;
  void* st;
};

int prove_impl(void* vp) {
  using clk = std::chrono::steady_clock;
  ProveCall& c = *static_cast<ProveCall*>(vp);
  const flk_blake3_setup& s = *c.s;
  void* st = c.st;
  const size_t m = s.m, nbl = s.n_blocks_log;
  const size_t n_f128 = size_t(1) << (m - 7);
  auto t0 = clk::now();
  double stages[6] = {0, 0, 0, 0, 0, 0};
  auto lap = [&](int i) {
    const auto t1 = clk::now();
    stages[i] = std::chrono::duration<double>(t1 - t0).count();
    t0 = t1;
  };

  // Witness (blake3.rs:4617-4647) and the PCS commit (CF/pcs/commit.rs:187-227).
  DeviceVec<F128> z(n_f128, st);
  DeviceVec<F128> a(n_f128, st), b(n_f128, st);
  DeviceVec<uint8_t> stripe(size_t(2048) << nbl, st);
  chk(flk_blake3_witness(c.blocks, c.n_blocks, nbl, z.data(), a.data(), b.data(), stripe.data(), st), "witness");
  // PcsParams { log_inv_rate: the profile's, log_batch_size: 6 } (blake3.rs:4721-4775)
  const size_t rate = s.ladder->levels[0].log_inv_rate;
  const size_t log_cols = m - 7 - 6, n_pos = size_t(1) << (log_cols + rate);
  DeviceVec<F128> codeword(n_pos << 6, st);
  DeviceVec<Hash> tree(2 * n_pos - 1, st);
  Hash root;
  chk(flk_commit(z.data(), log_cols, 6, rate, FLK_HASH_BLAKE3, codeword.data(), tree.data()->data(), root.data(),
                 st),
      "commit");
  lap(0);

  flock::challenger::FsChallenger fsc(std::string(reinterpret_cast<const char*>(c.domain), c.domain_len));
  Fs ch{fsc.raw()};
  // bind_statement (CF/proof.rs:51-59).
  ch.label("flock-r1cs-v0");
  ch.bytes(s.digest, 32);
  ch.bytes(root.data(), 32);

  ZcProof zp;
  ZcClaim zc;
  DeviceVec<F128> s_c(128, st);  // the c claim's s_hat_v, from zerocheck round 1
  zerocheck_prove(ch, reinterpret_cast<const uint8_t*>(a.data()), reinterpret_cast<const uint8_t*>(b.data()),
                  reinterpret_cast<const uint8_t*>(z.data()), m, zp, zc, st, s_c.data());
  a = DeviceVec<F128>();
  b = DeviceVec<F128>();
  lap(1);

  // x_ab = (z, mlv[..8], mlv[8..]) (BlockR1cs::x_ab_from_mlv, RowMajor; CF/r1cs.rs:241-270).
  std::vector<F128> x_ab;
  x_ab.push_back(zc.z);
  x_ab.insert(x_ab.end(), zc.mlv.begin(), zc.mlv.end());
  LcProof lp;
  LcClaim lc;
  lincheck_prove(ch, s.r1cs, stripe.data(), m, x_ab.data(), lp, lc, st);
  stripe = DeviceVec<uint8_t>();
  lap(2);

  // The two z-claims, address-ordered for the PCS (ab_claim_point / c_claim_point,
  // CF/r1cs.rs:272-330): x_full = x_inner_rest || x_outer.
  const size_t rest = K_LOG - K_SKIP;
  std::vector<F128> x_full_ab = lc.r_inner_rest;
  x_full_ab.insert(x_full_ab.end(), zc.mlv.begin() + rest, zc.mlv.end());
  std::vector<std::vector<F128>> x_fulls{x_full_ab, zc.r_rest};
  // pcs::compute_combined_basis_and_target (CF/pcs.rs:778-870); the c claim's
  // s_hat_v is the fold zerocheck round 1 made of the same bits against the same
  // eq (suffix r[7..m)), so the witness is folded once, for the ab claim.
  flock::RingSwitchBatch rs;
  {
    const size_t log_len = m - 7;
    std::vector<F128> xs, shv(256);
    for (const auto& x : x_fulls) xs.insert(xs.end(), x.begin(), x.end());
    const F128* known[2] = {nullptr, s_c.data()};
    rs.sumcheck_claims.resize(2);
    rs.gammas.resize(2);
    rs.b_combined = DeviceVec<F128>(z.len(), st);
    chk(flk::pcs_combined_basis_and_target_known(fsc.raw(), z.data(), log_len, xs.data(), 2, known, shv.data(),
                                                 rs.sumcheck_claims.data(), rs.gammas.data(), rs.b_combined.data(),
                                                 &rs.target_combined, st),
        "pcs_combined_basis_and_target");
    for (size_t j = 0; j < 2; ++j) rs.s_hat_v.emplace_back(shv.begin() + 128 * j, shv.begin() + 128 * (j + 1));
  }
  s_c = DeviceVec<F128>();
  lap(3);

  const flock::ligerito::ProverConfig cfg = prover_config(*s.ladder);
  flock::ligerito::LigeritoProof lig = flock::ligerito::recursive_prover_with_basis(
      cfg, std::move(z), std::move(rs.b_combined), rs.target_combined, codeword, tree, fsc, st);
  lap(4);

  Writer w;
  w.b.reserve(size_t(1) << 19);
  w.raw("FLOCK", 5);
  w.b.push_back(4);  // VERSION
  w.b.push_back(2);  // flavor: R1cs Ligerito
  // Commitment { root, PcsParams { m, log_inv_rate, log_batch_size, profile, merkle_hash } }
  w.raw(root.data(), 32);
  w.u64(m);
  w.u64(s.ladder->levels[0].log_inv_rate);
  w.u64(6);
  w.u32(uint32_t(s.profile));
  w.u32(1);  // HashKind::Blake3
  // ZerocheckProof
  w.f128_vec(zp.round1_ab);
  w.f128_vec(zp.round1_c);
  w.pairs(zp.rounds);
  w.f128(zp.final_a);
  w.f128(zp.final_b);
  w.f128(zp.final_c);
  // LincheckProof
  w.pairs(lp.rounds);
  w.f128_vec(lp.z_partial);
  // BatchOpeningProofLigerito { ring_switches, ligerito }
  w.u64(rs.s_hat_v.size());
  for (const auto& v : rs.s_hat_v) w.f128_vec(v);
  w.raw(lig.initial_root.data(), 32);
  w.rows(lig.initial_proof.rows, size_t(1) << cfg.initial_k);
  w.hashes(lig.initial_proof.merkle_proof);
  w.hashes(lig.recursive_roots);
  w.u64(lig.recursive_proofs.size());
  for (size_t i = 0; i < lig.recursive_proofs.size(); ++i) {
    w.rows(lig.recursive_proofs[i].rows, size_t(1) << cfg.recursive_ks[i]);
    w.hashes(lig.recursive_proofs[i].merkle_proof);
  }
  w.f128_vec(lig.yr);
  w.rows(lig.final_proof.rows, size_t(1) << cfg.recursive_ks.back());
  w.hashes(lig.final_proof.merkle_proof);
  w.u64(lig.sumcheck_transcript.size());
  for (const auto& msg : lig.sumcheck_transcript) {
    w.f128(msg.u_0);
    w.f128(msg.u_2);
  }
  w.u64_vec(lig.grinding_nonces);
  w.f128_vec(lig.ood_values);
  w.u64_vec(lig.fold_grinding_nonces);
  lap(5);

  if (c.out_len) *c.out_len = w.b.size();
  if (c.stage_s) std::copy(stages, stages + 6, c.stage_s);
  if (w.b.size() > c.cap)
    return flk::set_error(FLK_E_SHAPE, "blake3_prove: proof of %zu bytes exceeds the %zu-byte buffer", w.b.size(),
                          c.cap);
  std::memcpy(c.out, w.b.data(), w.b.size());
  return FLK_OK;
}

}  // namespace

extern "C" {

int flk_zerocheck_prove(flk_fs* fs, const uint8_t* a_bits, const uint8_t* b_bits, const uint8_t* c_bits, size_t m,
                        flk_f128* round1_ab, flk_f128* round1_c, flk_f128* rounds, flk_f128* finals,
                        flk_f128* z_out, flk_f128* mlv_out, flk_f128* r_rest_out, void* stream) {
  if (int rc = flk::check_enabled()) return rc;
  if (!fs || !a_bits || !b_bits || !c_bits || !round1_ab || !round1_c || !rounds || !finals || !z_out || !mlv_out ||
      !r_rest_out)
    return flk::set_error(FLK_E_SHAPE, "zerocheck_prove: null pointer");
  if (m < K_SKIP + N_INNER || m > 40) return flk::set_error(FLK_E_SHAPE, "zerocheck_prove: m=%zu outside 13..40", m);
  struct Ctx {
    flk_fs* fs;
    const uint8_t *a, *b, *c;
    size_t m;
    flk_f128 *r1ab, *r1c, *rounds, *finals, *z, *mlv, *rr;
    void* st;
  } ctx{fs, a_bits, b_bits, c_bits, m, round1_ab, round1_c, rounds, finals, z_out, mlv_out, r_rest_out, stream};
  return guarded("zerocheck_prove", [](void* p) -> int {
    Ctx& c = *static_cast<Ctx*>(p);
    Fs ch{c.fs};
    ZcProof pf;
    ZcClaim cl;
    zerocheck_prove(ch, c.a, c.b, c.c, c.m, pf, cl, c.st);
    std::copy(pf.round1_ab.begin(), pf.round1_ab.end(), c.r1ab);
    std::copy(pf.round1_c.begin(), pf.round1_c.end(), c.r1c);
    for (size_t i = 0; i < pf.rounds.size(); ++i) {
      c.rounds[2 * i] = pf.rounds[i][0];
      c.rounds[2 * i + 1] = pf.rounds[i][1];
    }
    c.finals[0] = pf.final_a;
    c.finals[1] = pf.final_b;
    c.finals[2] = pf.final_c;
    *c.z = cl.z;
    std::copy(cl.mlv.begin(), cl.mlv.end(), c.mlv);
    std::copy(cl.r_rest.begin(), cl.r_rest.end(), c.rr);
    return FLK_OK;
  }, &ctx);
}

int flk_lincheck_prove(flk_fs* fs, const flk_r1cs* r1cs, const uint8_t* z_stripe, const flk_f128* x_ab,
                       flk_f128* rounds, flk_f128* z_partial, flk_f128* r_inner_skip, flk_f128* r_inner_rest,
                       flk_f128* w, void* stream) {
  if (int rc = flk::check_enabled()) return rc;
  if (!fs || !r1cs || !z_stripe || !x_ab || !rounds || !z_partial || !r_inner_skip || !r_inner_rest || !w)
    return flk::set_error(FLK_E_SHAPE, "lincheck_prove: null pointer");
  size_t m = 0;
  if (int rc = flk_r1cs_shape(r1cs, &m, nullptr, nullptr, nullptr, nullptr)) return rc;
  struct Ctx {
    flk_fs* fs;
    const flk_r1cs* r;
    const uint8_t* z;
    const flk_f128* x;
    size_t m;
    flk_f128 *rounds, *zp, *rs, *rr, *w;
    void* st;
  } ctx{fs, r1cs, z_stripe, x_ab, m, rounds, z_partial, r_inner_skip, r_inner_rest, w, stream};
  return guarded("lincheck_prove", [](void* p) -> int {
    Ctx& c = *static_cast<Ctx*>(p);
    Fs ch{c.fs};
    LcProof pf;
    LcClaim cl;
    lincheck_prove(ch, c.r, c.z, c.m, c.x, pf, cl, c.st);
    for (size_t i = 0; i < pf.rounds.size(); ++i) {
      c.rounds[2 * i] = pf.rounds[i][0];
      c.rounds[2 * i + 1] = pf.rounds[i][1];
    }
    std::copy(pf.z_partial.begin(), pf.z_partial.end(), c.zp);
    *c.rs = cl.r_inner_skip;
    std::copy(cl.r_inner_rest.begin(), cl.r_inner_rest.end(), c.rr);
    *c.w = cl.w;
    return FLK_OK;
  }, &ctx);
}

int flk_blake3_statement_digest(size_t m, uint8_t* out) {
  if (!out) return flk::set_error(FLK_E_SHAPE, "blake3_statement_digest: null pointer");
  if (m < K_LOG + 3 || m > K_LOG + 26) return flk::set_error(FLK_E_SHAPE, "blake3_statement_digest: m=%zu", m);
  flk::r1cs::statement_digest(m, out);
  return FLK_OK;
}

}  // extern "C"

namespace flk {

size_t blake3_arena_bytes(const flk_blake3_setup* s) { return s->arena_bytes; }
size_t blake3_setup_m(const flk_blake3_setup* s) { return s->m; }

int blake3_prove_in(const flk_blake3_setup* s, Arena* arena, const flk_compression* blocks, size_t n_blocks,
                    const uint8_t* domain, size_t domain_len, uint8_t* proof_out, size_t cap, size_t* proof_len,
                    double* stage_seconds, void* stream) {
  if (int rc = flk::check_enabled()) return rc;
  if (!s || (n_blocks && !blocks) || (domain_len && !domain) || !proof_out)
    return flk::set_error(FLK_E_SHAPE, "blake3_prove: null pointer");
  if (n_blocks > (size_t(1) << s->n_blocks_log))
    return flk::set_error(FLK_E_SHAPE, "blake3_prove: %zu compressions for %zu slots", n_blocks,
                          size_t(1) << s->n_blocks_log);
  ProveCall c{s, blocks, n_blocks, domain, domain_len, proof_out, cap, proof_len, stage_seconds, stream};
  int rc;
  {
    ArenaScope scope(arena);
    rc = guarded("blake3_prove", prove_impl, &c);
  }
  if (arena) {
    // Nothing still queued may touch the arena when another prove takes it.
    const int sr = flk_synchronize(stream);
    if (rc == FLK_OK) rc = sr;
  }
  return rc;
}

}  // namespace flk

extern "C" {

int flk_blake3_setup_new(size_t n_blocks, int profile, flk_blake3_setup** out) {
  if (int rc = flk::check_enabled()) return rc;
  if (!out) return flk::set_error(FLK_E_SHAPE, "blake3_setup_new: null pointer");
  if (n_blocks < 1) return flk::set_error(FLK_E_SHAPE, "blake3_setup_new: n_blocks must be >= 1");
  size_t nbl = 3;  // min_n_blocks_log: ceil(log2(max(n, 8))) (blake3.rs:1854-1858)
  while ((size_t(1) << nbl) < n_blocks) ++nbl;
  const size_t m = K_LOG + nbl;
  const Ladder* l = find_ladder(m, profile);
  if (!l || l->initial_k != 6)
    return flk::set_error(FLK_E_SHAPE, "blake3_setup_new: no Ligerito ladder for m=%zu profile=%d (m = 22..35)", m,
                          profile);
  flk_r1cs* r = nullptr;
  if (int rc = flk_blake3_r1cs_new(nbl, &r)) return rc;
  flk_blake3_setup* s = new (std::nothrow) flk_blake3_setup();
  if (s) {
    s->n_blocks_log = nbl;
    s->m = m;
    s->profile = profile;
    s->ladder = l;
    s->r1cs = r;
  }
  if (!s) {
    flk_r1cs_free(r);
    return flk::set_error(FLK_E_NOMEM, "blake3_setup_new: out of host memory");
  }
  flk::r1cs::statement_digest(m, s->digest);
  // The measured peak is below the footprint estimate (6.51 GB against 7.32 GB at m = 32); 1/8 and
  // 64 MiB on top cover the estimate's rounding and the size-independent workspaces.
  const size_t fp = prove_footprint(m, l->levels[0].log_inv_rate);
  s->arena_bytes = fp + fp / 8 + (size_t(64) << 20);
  *out = s;
  return FLK_OK;
}

void flk_blake3_setup_free(flk_blake3_setup* s) {
  if (!s) return;
  for (flk::Arena* a : s->arenas) flk::arena_delete(a);
  flk_r1cs_free(s->r1cs);
  delete s;
}

int flk_blake3_setup_info(const flk_blake3_setup* s, size_t* m, size_t* n_blocks_log, size_t* device_bytes,
                          uint8_t* digest_out) {
  if (!s) return flk::set_error(FLK_E_SHAPE, "blake3_setup_info: null pointer");
  if (m) *m = s->m;
  if (n_blocks_log) *n_blocks_log = s->n_blocks_log;
  if (device_bytes) *device_bytes = prove_footprint(s->m, s->ladder->levels[0].log_inv_rate);
  if (digest_out) std::memcpy(digest_out, s->digest, 32);
  return FLK_OK;
}

int flk_blake3_proof_bound(const flk_blake3_setup* s, size_t* bytes) {
  if (!s || !bytes) return flk::set_error(FLK_E_SHAPE, "blake3_proof_bound: null pointer");
  const Ladder& l = *s->ladder;
  // Fixed part: header, commitment, zerocheck (2 x 64 + 2 (m - 6) + 3 values),
  // lincheck (16 + 64 values), two ring switches (2 x 128 values), length words.
  size_t b = 4096 + 16 * (128 + 2 * (s->m - 6) + 3 + 16 + 64 + 256);
  size_t msgs = 1 + l.initial_k, nonces = 0, queries = 0, row = 1, depth = 1;
  for (uint32_t i = 0; i < l.n_levels; ++i) {
    const Level& v = l.levels[i];
    row = std::max(row, size_t(1) << v.log_num_interleaved);
    depth = std::max(depth, size_t(v.log_msg_cols + v.log_inv_rate + 1));
    queries += v.queries;
    msgs += v.k_recursive + v.ood_samples + 1;
    nonces += 1 + v.fold_grinding_bits;
  }
  // opened rows (length word + row per query) and multi-proof hashes at every level
  b += 16 * l.n_levels + queries * (8 + 16 * row) + queries * depth * 32;
  b += 16 * (size_t(1) << l.yr_log_n) + 32 * msgs + 8 * nonces + 16 * 64 + 32 * (l.n_levels + 1);
  *bytes = b;
  return FLK_OK;
}

int flk_blake3_prove(const flk_blake3_setup* s, const flk_compression* blocks, size_t n_blocks, const uint8_t* domain,
                     size_t domain_len, uint8_t* proof_out, size_t cap, size_t* proof_len, double* stage_seconds,
                     void* stream) {
  if (int rc = flk::check_enabled()) return rc;
  if (!s || (n_blocks && !blocks) || (domain_len && !domain) || !proof_out)
    return flk::set_error(FLK_E_SHAPE, "blake3_prove: null pointer");
  if (n_blocks > (size_t(1) << s->n_blocks_log))
    return flk::set_error(FLK_E_SHAPE, "blake3_prove: %zu compressions for %zu slots", n_blocks,
                          size_t(1) << s->n_blocks_log);
  flk_blake3_setup* ms = const_cast<flk_blake3_setup*>(s);  // the arena list is the only mutable state
  flk::Arena* arena = nullptr;
  {
    std::lock_guard<std::mutex> lk(ms->arena_mu);
    if (!ms->idle.empty()) {
      arena = ms->idle.back();
      ms->idle.pop_back();
    }
  }
  if (!arena) {
    int st = FLK_OK;
    arena = flk::arena_new(s->arena_bytes, &st);  // on failure the prove uses the stream-ordered pool
    if (arena) {
      std::lock_guard<std::mutex> lk(ms->arena_mu);
      ms->arenas.push_back(arena);
    }
  }
  const int rc = flk::blake3_prove_in(s, arena, blocks, n_blocks, domain, domain_len, proof_out, cap, proof_len,
                                      stage_seconds, stream);
  if (arena) {
    std::lock_guard<std::mutex> lk(ms->arena_mu);
    ms->idle.push_back(arena);
  }
  return rc;
}

int flk_blake3_setup_arena_stats(const flk_blake3_setup* s, size_t* capacity, size_t* peak, size_t* n_arenas,
                                 size_t* fallbacks) {
  if (!s) return flk::set_error(FLK_E_SHAPE, "blake3_setup_arena_stats: null pointer");
  flk_blake3_setup* ms = const_cast<flk_blake3_setup*>(s);
  std::lock_guard<std::mutex> lk(ms->arena_mu);
  size_t pk = 0, fb = 0;
  for (flk::Arena* a : ms->arenas) {
    size_t p = 0, f = 0;
    flk::arena_stats(a, nullptr, nullptr, &p, &f);
    pk = std::max(pk, p);
    fb += f;
  }
  if (capacity) *capacity = s->arena_bytes;
  if (peak) *peak = pk;
  if (n_arenas) *n_arenas = ms->arenas.size();
  if (fallbacks) *fallbacks = fb;
  return FLK_OK;
}

}  // extern "C"
