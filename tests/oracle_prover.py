"""Test infrastructure: a CPU prover for the BLAKE3 R1CS proof, sequenced over
the oracle's C restatements (oracle/*.c through oracle_ffi) in the reference's
prover order (CF = crates/flock-core/src/, CP = crates/flock-prover/src/):

  witness             r1cs_hashes::blake3 witness (CP/r1cs_hashes/blake3.rs:4617-4647)
  commit              pcs::commit (CF/pcs/commit.rs:187-227): log_batch_size 6, rate 1/2
  bind_statement      CF/proof.rs:51-59
  zerocheck           zerocheck::prove_packed_padded (CF/zerocheck.rs:384-535):
                      round1_naive, P^C(z) on Lambda, fold at z, the multilinear rounds
                      (CF/zerocheck/multilinear.rs:575-640, 3194-3213)
  lincheck            lincheck::prove_padded (CF/lincheck.rs:5345-5878)
  PCS                 compute_combined_basis_and_target (CF/pcs.rs:778-870)
  Ligerito            recursive_prover_with_basis (CF/pcs/ligerito.rs:10185-11040)
  serialisation       R1csProofBundleLigerito::to_bytes (CP/proof_io.rs:160-198)

It shares nothing with the device library: every field product, hash, fold and
grind is the oracle's.  Two uses: the device proof is compared with its bytes,
and its wall time on the host's cores is the CPU baseline of the proof bench.
"""
import time

import numpy as np

import flock_verifier as fv
from oracle_ffi import F128_DT, f128_to_ints, OracleFs

K_LOG, K_SKIP, CONST_POS = fv.K_LOG, fv.K_SKIP, fv.CONST_POS
STAGES = ("witness_commit", "zerocheck", "lincheck", "ring_switch", "ligerito", "serialize")


def _first(a):
    return f128_to_ints(a[:1])[0]


def _pair(msg):
    v = f128_to_ints(msg)
    return v[0], v[1]


def zerocheck_prove(orc, ch, a_bytes, b_bytes, c_bytes, m, threads=1):
    n_mlv = m - K_SKIP
    ch.observe_label(b"flock-zerocheck-v0")
    r_skip = ch.sample_f128_vec(K_SKIP)
    r_outer = ch.sample_f128_vec(m - K_SKIP - 7)
    r = r_skip + fv.fixed_challenges(orc) + r_outer
    if threads > 1:
        pab, pc = orc.zc_round1_mt(a_bytes, b_bytes, c_bytes, m, K_SKIP, r, threads)
    else:
        pab, pc = orc.zc_round1(a_bytes, b_bytes, c_bytes, m, K_SKIP, r)
    pf = {"round1_ab": f128_to_ints(pab), "round1_c": f128_to_ints(pc), "rounds": []}
    ch.observe_f128_slice(pf["round1_ab"])
    ch.observe_f128_slice(pf["round1_c"])
    z = ch.sample_f128()
    phi = fv.phi8_nodes(orc)
    pf["final_c"] = fv.inner(orc, fv.lagrange_on_nodes(orc, phi[64:128], z), pf["round1_c"])
    w = orc.zc_lagrange_weights(K_SKIP, z)
    am = orc.zc_fold_at_z(a_bytes, m, K_SKIP, w)
    bm = orc.zc_fold_at_z(b_bytes, m, K_SKIP, w)
    rhos = []
    for i in range(n_mlv):
        if i:
            am = orc.partial_eval_lsb(am, rhos[-1:])
            bm = orc.partial_eval_lsb(bm, rhos[-1:])
        g = _pair(orc.zc_round_pair(am, bm, orc.build_eq(r[K_SKIP + i + 1:])))
        pf["rounds"].append(g)
        ch.observe_f128(g[0])
        ch.observe_f128(g[1])
        rhos.append(ch.sample_f128())
    pf["final_a"] = _first(orc.partial_eval_lsb(am, rhos[-1:]))
    pf["final_b"] = _first(orc.partial_eval_lsb(bm, rhos[-1:]))
    ch.observe_f128(pf["final_a"])
    ch.observe_f128(pf["final_b"])
    return pf, {"z": z, "mlv": rhos, "r_rest": r[K_SKIP:]}


def lincheck_prove(orc, ch, stripe, m, z_skip, mlv):
    rest = K_LOG - K_SKIP
    ch.observe_label(b"flock-lincheck-v0")
    alpha = ch.sample_f128()
    beta = ch.sample_f128()
    comb = fv.lincheck_comb(orc, alpha, fv.quirky_eq(orc, z_skip, mlv[:rest]))
    cp = _first(comb[CONST_POS:CONST_POS + 1]) ^ beta
    comb[CONST_POS] = (cp & (2**64 - 1), cp >> 64)
    zv = orc.lc_partial_fold(stripe, m, K_LOG, orc.build_eq(mlv[rest:]))
    pf, rr = {"rounds": []}, []
    for _ in range(rest):
        q = _pair(orc.lc_round_msg_msb(comb, zv))
        ch.observe_f128(q[0])
        ch.observe_f128(q[1])
        ri = ch.sample_f128()
        pf["rounds"].append(q)
        rr.append(ri)
        comb = orc.lc_bind_top(comb, ri)
        zv = orc.lc_bind_top(zv, ri)
    pf["z_partial"] = f128_to_ints(zv)
    ch.observe_f128_slice(pf["z_partial"])
    r_skip = ch.sample_f128()
    w = fv.inner(orc, f128_to_ints(orc.zc_lagrange_weights(K_SKIP, r_skip)), pf["z_partial"])
    return pf, {"r_inner_skip": r_skip, "r_inner_rest": rr[::-1], "w": w}


def _grind(orc, ch, bits):
    """grind_pow with the smallest valid nonce, absorbed as u64le (CF/challenger.rs:549-601)."""
    nonce = orc.pow_grind(ch.state_digest(), bits)
    ch.observe_bytes(int(nonce).to_bytes(8, "little"))
    return nonce


def _open(orc, cw, tree, row_len, queries):
    n = len(cw) // row_len
    rows = [cw[q * row_len:(q + 1) * row_len].copy() for q in queries]
    return rows, [bytes(tree[i]) for i in orc.multi_proof_indices(n, queries)]


def ligerito_prove(orc, cfg, ch, f, b, target, l0_cw, l0_tree, l0_root):
    k0, ks, rates = cfg["initial_k"], cfg["recursive_ks"], cfg["log_inv_rates"]
    r = len(ks)
    log_n = int(len(f)).bit_length() - 1
    n_cur = log_n - k0
    block0 = 1 << (n_cur + rates[0])
    lig = {"initial_root": l0_root, "recursive_roots": [], "recursive_proofs": [], "sumcheck_transcript": [],
           "grinding_nonces": [], "ood_values": [], "fold_grinding_nonces": []}
    tx = lig["sumcheck_transcript"]

    def emit(msg):
        m = _pair(msg)
        tx.append(m)
        ch.observe_f128(m[0])
        ch.observe_f128(m[1])

    ch.observe_label(b"flock-ligerito-basis-v0")
    ch.observe_f128(target)
    ch.observe_bytes(l0_root)
    emit(orc.round_msg(f, b))

    def fold_rounds(k, level):
        nonlocal f, b
        rs = []
        for j in range(k):
            fb = cfg["fold_grinding_bits"][level]
            if fb > j:
                lig["fold_grinding_nonces"].append(_grind(orc, ch, fb - j))
            ri = ch.sample_f128()
            rs.append(ri)
            f, b, msg = orc.fold_and_msg(f, b, ri)
            emit(msg)
        return rs

    def commit(k_next, rate):
        cw, tree, root = orc.commit(f, 1 << k_next, rate)
        lig["recursive_roots"].append(root)
        ch.observe_bytes(root)
        return cw, tree

    def ood(level):
        nonlocal b
        b, _, ys, msgs = orc.ligerito_ood(ch, f, b, cfg["ood_samples"][level], 0)
        lig["ood_values"] += ys
        tx.extend(msgs)

    def introduce(cols, rows, v, q, alpha):
        nonlocal b
        basis, _ = orc.induce(cols, np.array(rows), v, q, alpha)
        emit(orc.round_msg(f, basis))
        b = orc.glue(b, basis, ch.sample_f128())

    r_lane = fold_rounds(k0, 0)
    prev_cw, prev_tree = commit(ks[0], rates[1])
    prev_k, prev_block = ks[0], 1 << (n_cur - ks[0] + rates[1])
    ood(1)
    lig["grinding_nonces"].append(_grind(orc, ch, cfg["grinding_bits"][0]))
    q0 = ch.sample_distinct_queries(block0, cfg["queries"][0])
    a0 = ch.sample_f128_vec(fv.ceil_log2(cfg["queries"][0]))
    lig["initial_proof"] = _open(orc, l0_cw, l0_tree, 1 << k0, q0)
    introduce(n_cur, lig["initial_proof"][0], r_lane, q0, a0)
    for i in range(r):
        level_rs = fold_rounds(ks[i], i + 1)
        n_cur -= ks[i]
        if i == r - 1:
            lig["yr"] = f128_to_ints(f)
            for v in lig["yr"]:
                ch.observe_f128(v)
            lig["grinding_nonces"].append(_grind(orc, ch, cfg["grinding_bits"][i + 1]))
            ql = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
            lig["final_proof"] = _open(orc, prev_cw, prev_tree, 1 << prev_k, ql)
            return lig
        nxt = commit(ks[i + 1], rates[i + 2])
        ood(i + 2)
        lig["grinding_nonces"].append(_grind(orc, ch, cfg["grinding_bits"][i + 1]))
        qi = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
        ai = ch.sample_f128_vec(fv.ceil_log2(cfg["queries"][i + 1]))
        lig["recursive_proofs"].append(_open(orc, prev_cw, prev_tree, 1 << prev_k, qi))
        introduce(n_cur, lig["recursive_proofs"][-1][0], level_rs, qi, ai)
        prev_cw, prev_tree = nxt
        prev_k, prev_block = ks[i + 1], 1 << (n_cur - ks[i + 1] + rates[i + 2])
    raise AssertionError("unreachable")


def prove(orc, comps, n_blocks_log, domain=fv.DOMAIN, threads=8, digest=None, profile=0):
    """The proof bytes of `comps` (COMPRESSION_DT) padded to 2^n_blocks_log blocks, and
    the wall time of each stage.  `digest` is the statement digest of m (computed when
    None: it is per-m set-up, like the device's, and outside the stage times)."""
    m = K_LOG + n_blocks_log
    cfg = fv.ladders()[(m, profile)]
    if digest is None:
        digest = fv.statement_digest(orc, m)
    t = [time.perf_counter()]
    lap = lambda: t.append(time.perf_counter())  # noqa: E731
    z, a, bw, stripe = orc.witness_mt(comps, n_blocks_log, threads)
    zf = z.view(F128_DT)
    l0_cw, l0_tree, root = orc.commit_mt(zf, 64, cfg["log_inv_rates"][0], threads)
    lap()
    ch = OracleFs(orc, domain)
    ch.observe_label(b"flock-r1cs-v0")
    ch.observe_bytes(digest)
    ch.observe_bytes(root)
    as_bytes = lambda w: w.view(np.uint8)  # noqa: E731
    zp, zc = zerocheck_prove(orc, ch, as_bytes(a), as_bytes(bw), as_bytes(z), m, threads)
    del a, bw
    lap()
    lp, lc = lincheck_prove(orc, ch, stripe, m, zc["z"], zc["mlv"])
    del stripe
    lap()
    rest = K_LOG - K_SKIP
    x_fulls = [lc["r_inner_rest"] + zc["mlv"][rest:], zc["r_rest"]]
    s_hat_vs, _, _, b_comb, target = orc.pcs_combined_basis_and_target(ch, zf, x_fulls)
    lap()
    lig = ligerito_prove(orc, cfg, ch, zf.copy(), b_comb, target, l0_cw, l0_tree, root)
    lap()
    bundle = {"root": root, "m": m, "log_inv_rate": cfg["log_inv_rates"][0], "log_batch_size": K_SKIP,
              "profile": profile, "merkle_hash": 1, "zc": zp, "lc": lp, "s_hat_v": s_hat_vs, "lig": lig}
    out = fv.serialize_bundle(bundle)
    lap()
    return out, dict(zip(STAGES, np.diff(t).tolist()))
