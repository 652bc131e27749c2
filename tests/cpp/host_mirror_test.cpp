// Parity of the C++ host mirror (include/flock_host.hpp) on the GPU against the
// oracle, written the way the reference's own tests read
// (CF/pcs/commit.rs:1838-1877 commit_matches_full_ntt_oracle; the SumcheckProver
// fold/introduce/glue sequence of CF/pcs/ligerito.rs:8906-9275).  Exit code 0 = pass.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "flock_host.hpp"
#include "flock_oracle.h"

using flock::F128;
using flock::DeviceVec;
using flock::operator+;
using flock::operator*;
using flock::operator==;

static std::vector<F128> rand_f128(uint64_t seed, size_t n) {
  std::vector<F128> v(n);
  orc_fill_f128(seed, reinterpret_cast<orc_f128*>(v.data()), n);
  return v;
}
static bool same(const std::vector<F128>& a, const std::vector<F128>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (!(a[i] == b[i])) return false;
  return true;
}
#define EXPECT(cond, what)                      \
  do {                                          \
    if (!(cond)) {                              \
      std::fprintf(stderr, "FAIL: %s\n", what); \
      return 1;                                 \
    }                                           \
  } while (0)

static orc_f128 O(F128 a) { return orc_f128{a.lo, a.hi}; }

static void dump_f128s(FILE* o, const char* key, const std::vector<F128>& v) {
  std::fprintf(o, "%s", key);
  for (const F128& x : v)
    std::fprintf(o, " %016llx%016llx", (unsigned long long)x.hi, (unsigned long long)x.lo);
  std::fprintf(o, "\n");
}
static void dump_hashes(FILE* o, const char* key, const std::vector<flock::Hash>& v) {
  std::fprintf(o, "%s", key);
  for (const auto& h : v) {
    std::fprintf(o, " ");
    for (uint8_t c : h) std::fprintf(o, "%02x", c);
  }
  std::fprintf(o, "\n");
}
static void dump_u64s(FILE* o, const char* key, const std::vector<uint64_t>& v) {
  std::fprintf(o, "%s", key);
  for (uint64_t x : v) std::fprintf(o, " %llu", (unsigned long long)x);
  std::fprintf(o, "\n");
}

// The recursive Ligerito open with the m = 32 Fast ladder on a 2^25-element
// (f, b) pair and its L0 commitment: prints the median wall time of 5 opens.
static int bench_open_m32() {
  const size_t log_n = 25, k0 = 6, n = size_t(1) << log_n;
  DeviceVec<F128> f(rand_f128(0xF00D, n)), b(rand_f128(0xBEEF, n));
  auto l0 = flock::pcs::ligero_commit(f, log_n - k0, k0, 1, flock::merkle::HashKind::Blake3);
  DeviceVec<F128> out3(3);
  flock::check(flk_sc_round_msg_and_eval(f.data(), b.data(), n, out3.data(), nullptr), "round_msg_and_eval");
  const F128 target = out3.to_host()[2];
  flock::ligerito::ProverConfig cfg{k0, {3, 3, 3, 3, 3}, {1, 2, 3, 4, 5, 6}, {218, 106, 71, 53, 43, 36},
                                    {0, 1, 1, 1, 1, 1}, {0, 0, 0, 0, 0, 0}, {19, 14, 11, 8, 6, 4}};
  std::vector<double> ms;
  for (int it = 0; it < 6; ++it) {
    DeviceVec<F128> fc(n), bc(n);
    flock::check(flk_copy_device_to_device(fc.data(), f.data(), n * 16, nullptr), "copy");
    flock::check(flk_copy_device_to_device(bc.data(), b.data(), n * 16, nullptr), "copy");
    flock::check(flk_synchronize(nullptr), "sync");
    flock::challenger::FsChallenger t("flock-bench");
    const auto t0 = std::chrono::steady_clock::now();
    auto p = flock::ligerito::recursive_prover_with_basis(cfg, std::move(fc), std::move(bc), target, l0.codeword,
                                                          l0.merkle_tree, t);
    flock::check(flk_synchronize(nullptr), "sync");
    const double dt = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (it > 0) ms.push_back(dt);  // the first open warms the twiddle caches
    if (p.yr.size() != 16) return 1;
    std::printf("open %d: %.3f ms\n", it, dt);
    std::fflush(stdout);
  }
  std::sort(ms.begin(), ms.end());
  std::printf("ligerito open m32 (C++ host, median of %zu): %.3f ms\n", ms.size(), ms[ms.size() / 2]);
  return 0;
}

int main(int argc, char** argv) {
  if (argc > 1 && std::strcmp(argv[1], "--bench-open-m32") == 0) return bench_open_m32();
  if (!flk_device_available()) {
    std::fprintf(stderr, "no device: %s\n", flk_last_error());
    return 2;
  }
  // commit_matches_full_ntt_oracle at (m, r, log_batch) = (14, 2, 3) and (18, 1, 6)
  const size_t shapes[2][3] = {{14, 2, 3}, {18, 1, 6}};
  for (auto& sh : shapes) {
    flock::pcs::PcsParams params{sh[0], sh[1], sh[2]};
    std::vector<F128> z = rand_f128(0xFEED + sh[0], size_t(1) << params.log_msg_len());
    DeviceVec<F128> zd(z);
    auto [commitment, pd] = flock::pcs::commit(zd, params);
    std::vector<F128> cw(params.codeword_len_f128());
    std::vector<uint8_t> tree(32 * (2 * params.n_leaves() - 1));
    uint8_t root[32];
    EXPECT(orc_commit(reinterpret_cast<const orc_f128*>(z.data()), z.size(), params.num_ntts(), params.log_inv_rate,
                      reinterpret_cast<orc_f128*>(cw.data()), tree.data(), root) == 0,
           "oracle commit");
    EXPECT(same(pd.codeword.to_host(), cw), "codeword == zero-padded full NTT");
    EXPECT(std::memcmp(commitment.root.data(), root, 32) == 0, "root == merkle_tree(codeword)");
    auto ntt = flock::ntt::AdditiveNttF128::standard(params.k_code());
    EXPECT(ntt.twiddle(0, 0) == (F128{0, 0}), "twiddle(0,0) == 0");
  }
  // SumcheckProver: new -> fold x2 -> introduce_new -> glue -> fold
  {
    const size_t n = 1 << 12;
    std::vector<F128> f = rand_f128(1, n), b = rand_f128(2, n);
    F128 h1{0, 0};
    for (size_t i = 0; i < n; ++i) h1 = h1 + f[i] * b[i];
    auto [prover, msg] = flock::ligerito::SumcheckProver::new_(DeviceVec<F128>(f), DeviceVec<F128>(b), h1);
    orc_f128 want[2];
    orc_sc_round_msg(reinterpret_cast<orc_f128*>(f.data()), reinterpret_cast<orc_f128*>(b.data()), n, want);
    EXPECT(msg.u_0.lo == want[0].lo && msg.u_2.hi == want[1].hi, "first message");
    std::vector<F128> cf = f, cb = b;
    uint64_t st = 7;
    for (int rnd = 0; rnd < 3; ++rnd) {
      if (rnd == 2) {
        std::vector<F128> bn = rand_f128(3, cf.size());
        F128 hn{orc_splitmix64_next(&st), orc_splitmix64_next(&st)};
        auto m2 = prover.introduce_new(DeviceVec<F128>(bn), hn);
        orc_sc_round_msg(reinterpret_cast<orc_f128*>(cf.data()), reinterpret_cast<orc_f128*>(bn.data()), cf.size(), want);
        EXPECT(m2.u_0.lo == want[0].lo && m2.u_2.lo == want[1].lo, "introduce_new message");
        F128 alpha{orc_splitmix64_next(&st), orc_splitmix64_next(&st)};
        prover.glue(alpha);
        orc_sc_glue(reinterpret_cast<orc_f128*>(cb.data()), reinterpret_cast<orc_f128*>(bn.data()), cb.size(), O(alpha));
        EXPECT(same(prover.combined_basis().to_host(), cb), "glued basis");
        EXPECT(prover.target() == h1 + alpha * hn, "glued target");
      }
      F128 r{orc_splitmix64_next(&st), orc_splitmix64_next(&st)};
      auto m = prover.fold(r);
      std::vector<F128> nf(cf.size() / 2), nb(cb.size() / 2);
      orc_sc_fold_and_msg(reinterpret_cast<orc_f128*>(cf.data()), reinterpret_cast<orc_f128*>(cb.data()), cf.size(),
                          O(r), reinterpret_cast<orc_f128*>(nf.data()), reinterpret_cast<orc_f128*>(nb.data()), want);
      cf = nf;
      cb = nb;
      EXPECT(m.u_0.lo == want[0].lo && m.u_0.hi == want[0].hi && m.u_2.lo == want[1].lo, "fold message");
    }
    EXPECT(same(prover.f().to_host(), cf), "folded f");
  }
  // ring switch fold_1b_rows
  {
    std::vector<F128> z = rand_f128(11, 5000), e = rand_f128(12, 5000), want(128);
    auto got = flock::ring_switch::fold_1b_rows(DeviceVec<F128>(z), DeviceVec<F128>(e)).to_host();
    orc_rs_fold_1b_rows(reinterpret_cast<orc_f128*>(z.data()), reinterpret_cast<orc_f128*>(e.data()), z.size(),
                        reinterpret_cast<orc_f128*>(want.data()));
    EXPECT(same(got, want), "fold_1b_rows");
  }
  // BLAKE3 witness: 13 compressions in 16 slots (3 padding), with the stripe
  {
    std::vector<flk_compression> comps(13);
    uint64_t st = 99;
    for (auto& c : comps) {
      uint32_t* w = reinterpret_cast<uint32_t*>(&c);
      for (int i = 0; i < 28; ++i) w[i] = (uint32_t)orc_splitmix64_next(&st);
    }
    auto wit = flock::r1cs_hashes::blake3::generate_witness_with_ab_packed_and_lincheck(
        DeviceVec<flk_compression>(comps), 4);
    std::vector<uint64_t> z(16 * 256), a(16 * 256), b(16 * 256);
    std::vector<uint8_t> stripe(16 * 2048);
    EXPECT(orc_blake3_witness(reinterpret_cast<const orc_compression*>(comps.data()), comps.size(), 4, z.data(),
                              a.data(), b.data(), stripe.data()) == 0,
           "oracle witness");
    auto gz = wit.z.to_host(), ga = wit.a.to_host(), gb = wit.b.to_host();
    EXPECT(std::memcmp(gz.data(), z.data(), z.size() * 8) == 0, "witness z");
    EXPECT(std::memcmp(ga.data(), a.data(), a.size() * 8) == 0, "witness a");
    EXPECT(std::memcmp(gb.data(), b.data(), b.size() * 8) == 0, "witness b");
    EXPECT(wit.z_lincheck.to_host() == stripe, "witness lincheck stripe");
  }
  // FsChallenger: device grind, host verify, challenges in lockstep
  {
    auto script = [] {
      flock::challenger::FsChallenger t("pow-test");
      t.observe_label("flock-pow-test");
      const uint8_t root[4] = {'r', 'o', 'o', 't'};
      t.observe_bytes(root, 4);
      return t;
    };
    auto prover = script();
    auto verifier = script();
    auto d = prover.state_digest();
    const uint64_t n = prover.grind_pow(12);
    EXPECT(n == orc_pow_grind(d.data(), 12), "grind_pow nonce == oracle");
    EXPECT(verifier.verify_pow(n, 12), "verify_pow accepts");
    EXPECT(prover.sample_f128() == verifier.sample_f128(), "challenges agree after the PoW");
  }
  // query openings: pow_scan, row gather and multi-proof on a committed codeword
  {
    std::array<uint8_t, 32> d{};
    for (int i = 0; i < 32; ++i) d[i] = (uint8_t)(3 * i + 1);
    uint64_t n = 0;
    EXPECT(flock::challenger::pow_scan(d, 0, 1 << 20, 9, &n), "pow_scan finds a nonce");
    EXPECT(n == orc_pow_grind(d.data(), 9), "pow_scan == the oracle's smallest nonce");
    flock::pcs::PcsParams params{18, 1, 3};  // 512 positions x 8 lanes
    std::vector<F128> z = rand_f128(0xABCD, size_t(1) << params.log_msg_len());
    auto [commitment, pd] = flock::pcs::commit(DeviceVec<F128>(z), params);
    const std::vector<uint64_t> q = {5, 6, 7, 100, 300, 511, 6};
    auto rows = flock::pcs::gather_rows(pd.codeword, params.num_ntts(), q).to_host();
    auto cw = pd.codeword.to_host();
    bool rows_ok = true;
    for (size_t i = 0; i < q.size(); ++i)
      for (size_t j = 0; j < params.num_ntts(); ++j)
        rows_ok = rows_ok && rows[i * params.num_ntts() + j] == cw[q[i] * params.num_ntts() + j];
    EXPECT(rows_ok, "gather_rows == codeword rows");
    auto proof = flock::merkle::merkle_multi_proof(pd.merkle_tree, params.n_leaves(), q);
    std::vector<uint64_t> idx(q.size() * 20);
    const size_t k = orc_merkle_multi_proof_indices(params.n_leaves(), q.data(), q.size(), idx.data());
    auto tree = pd.merkle_tree.to_host();
    bool proof_ok = proof.size() == k;
    for (size_t i = 0; proof_ok && i < k; ++i) proof_ok = proof[i] == tree[idx[i]];
    EXPECT(proof_ok, "merkle_multi_proof == the oracle's sibling hashes");
  }
  // ring switch of three claims and the combined basis and target
  {
    const size_t log_len = 9;
    std::vector<F128> z = rand_f128(0x5151, size_t(1) << log_len);
    std::vector<std::vector<F128>> xs;
    for (int k = 0; k < 3; ++k) xs.push_back(rand_f128(0x77 + k, log_len + 1));
    flock::challenger::FsChallenger t("rs-test");
    orc_fs* ot = orc_fs_new(reinterpret_cast<const uint8_t*>("rs-test"), 7);
    auto got = flock::pcs::compute_combined_basis_and_target(DeviceVec<F128>(z), xs, t);
    std::vector<orc_f128> flat, shv(3 * 128), cl(3), ga(3), b(z.size());
    for (const auto& x : xs)
      for (const auto& v : x) flat.push_back(orc_f128{v.lo, v.hi});
    orc_f128 tgt{0, 0};
    EXPECT(orc_pcs_combined_basis_and_target(ot, reinterpret_cast<const orc_f128*>(z.data()), log_len, flat.data(), 3,
                                             shv.data(), cl.data(), ga.data(), b.data(), &tgt) == 0,
           "oracle combine");
    bool ok = got.target_combined.lo == tgt.lo && got.target_combined.hi == tgt.hi;
    for (size_t k = 0; k < 3; ++k) {
      ok = ok && got.sumcheck_claims[k].lo == cl[k].lo && got.sumcheck_claims[k].hi == cl[k].hi;
      ok = ok && got.gammas[k].lo == ga[k].lo && got.gammas[k].hi == ga[k].hi;
      ok = ok && std::memcmp(got.s_hat_v[k].data(), shv.data() + 128 * k, 128 * 16) == 0;
    }
    auto gb = got.b_combined.to_host();
    ok = ok && std::memcmp(gb.data(), b.data(), b.size() * 16) == 0;
    uint8_t od[32];
    orc_fs_digest(ot, od);
    ok = ok && std::memcmp(t.state_digest().data(), od, 32) == 0;
    orc_fs_free(ot);
    EXPECT(ok, "compute_combined_basis_and_target == oracle (claims, gammas, s_hat_v, basis, target, transcript)");
  }
  // OOD samples on a sumcheck pair
  {
    const size_t log_n = 10;
    std::vector<F128> f = rand_f128(0x0D0D, size_t(1) << log_n), b = rand_f128(0x0E0E, size_t(1) << log_n);
    auto [sp, m0] = flock::ligerito::SumcheckProver::new_(DeviceVec<F128>(f), DeviceVec<F128>(b), F128{7, 9});
    (void)m0;
    flock::challenger::FsChallenger t("ood-test");
    orc_fs* ot = orc_fs_new(reinterpret_cast<const uint8_t*>("ood-test"), 8);
    auto ys = sp.ood(t, 2);
    std::vector<orc_f128> ob(b.size()), oys(2), oms(4);
    std::memcpy(ob.data(), b.data(), b.size() * 16);
    orc_f128 ot_r{7, 9};
    EXPECT(orc_ligerito_ood(ot, reinterpret_cast<const orc_f128*>(f.data()), ob.data(), log_n, 2, &ot_r, oys.data(),
                            oms.data()) == 0,
           "oracle ood");
    auto gb = sp.combined_basis().to_host();
    bool ok = std::memcmp(gb.data(), ob.data(), ob.size() * 16) == 0 && sp.target().lo == ot_r.lo &&
              sp.target().hi == ot_r.hi && std::memcmp(ys.data(), oys.data(), 2 * 16) == 0;
    const auto& tr = sp.transcript();
    ok = ok && tr.size() == 3 && tr[1].u_0.lo == oms[0].lo && tr[2].u_2.hi == oms[3].hi;
    uint8_t od[32];
    orc_fs_digest(ot, od);
    ok = ok && std::memcmp(t.state_digest().data(), od, 32) == 0;
    orc_fs_free(ot);
    EXPECT(ok, "SumcheckProver::ood == oracle (values, messages, basis, target, transcript)");
  }
  // recursive Ligerito prover (two recursive levels); with a path argument the
  // proof is written out for the Python-side verifier and prover comparison
  {
    const size_t log_n = 12, k0 = 3;
    std::vector<F128> f = rand_f128(7, size_t(1) << log_n), b = rand_f128(8, size_t(1) << log_n);
    F128 target{0, 0};
    for (size_t i = 0; i < f.size(); ++i) target = target + f[i] * b[i];
    auto l0 = flock::pcs::ligero_commit(DeviceVec<F128>(f), log_n - k0, k0, 1, flock::merkle::HashKind::Blake3);
    flock::ligerito::ProverConfig cfg{k0, {2, 2}, {1, 2, 2}, {8, 6, 5}, {0, 1, 1}, {2, 2, 2}, {3, 2, 1}};
    flock::challenger::FsChallenger t("flock-test-v0");
    auto p = flock::ligerito::recursive_prover_with_basis(cfg, DeviceVec<F128>(f), DeviceVec<F128>(b), target,
                                                          l0.codeword, l0.merkle_tree, t);
    EXPECT(p.yr.size() == 32 && p.recursive_roots.size() == 2 && p.recursive_proofs.size() == 1 &&
               p.grinding_nonces.size() == 3 && p.ood_values.size() == 2 && p.sumcheck_transcript.size() == 12,
           "ligerito proof shape");
    if (argc > 1) {
      FILE* o = std::fopen(argv[1], "w");
      EXPECT(o, "open the proof dump");
      dump_hashes(o, "initial_root", {p.initial_root});
      dump_f128s(o, "initial_rows", p.initial_proof.rows);
      dump_hashes(o, "initial_merkle", p.initial_proof.merkle_proof);
      dump_hashes(o, "recursive_roots", p.recursive_roots);
      dump_f128s(o, "recursive_rows", p.recursive_proofs[0].rows);
      dump_hashes(o, "recursive_merkle", p.recursive_proofs[0].merkle_proof);
      dump_f128s(o, "yr", p.yr);
      dump_f128s(o, "final_rows", p.final_proof.rows);
      dump_hashes(o, "final_merkle", p.final_proof.merkle_proof);
      std::vector<F128> msgs;
      for (const auto& m : p.sumcheck_transcript) {
        msgs.push_back(m.u_0);
        msgs.push_back(m.u_2);
      }
      dump_f128s(o, "sumcheck_transcript", msgs);
      dump_u64s(o, "grinding_nonces", p.grinding_nonces);
      dump_u64s(o, "fold_grinding_nonces", p.fold_grinding_nonces);
      dump_f128s(o, "ood_values", p.ood_values);
      std::fclose(o);
    }
  }
  std::printf("host mirror parity ok\n");
  return 0;
}
