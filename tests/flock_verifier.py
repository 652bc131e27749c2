"""Test infrastructure: the reference's verifier for a serialised BLAKE3 R1CS
proof, restated over the oracle's primitives (CF = crates/flock-core/src/,
CP = crates/flock-prover/src/, BT = benchmark-tools/).

  parse_bundle                 R1csProofBundleLigerito::from_bytes with the harness's
                               options (CP/proof_io.rs:160-216, BT/harness/src/main.rs:393-407):
                               "FLOCK", version 4, flavor 2, bincode fixint, 500,000-byte
                               limit, no trailing bytes
  statement_digest             BlockR1cs::statement_digest (CF/r1cs.rs:334-385) over the
                               symbolic BLAKE3 matrices (test_oracle_witness.build_matrices)
  zerocheck_verify             zerocheck::verify (CF/zerocheck.rs:1642-1800)
  lincheck_verify              lincheck::verify (CF/lincheck.rs:5880-6040)
  ring_switch_verify_succinct  ring_switch::verify_succinct (CF/pcs/ring_switch.rs:4783-4815)
  eval_rs_eq_prefix / _finish  (:4864-4960) over TensorAlgebra (CF/pcs/tensor_algebra.rs)
  ligerito_verify_succinct     recursive_verifier_with_basis_succinct
                               (CF/pcs/ligerito.rs:11321-11834)
  verify_pcs                   verify_opening_batch_ligerito_mixed (CF/pcs.rs:1861-1998)
  verify_r1cs                  verify_ligerito / verify_core (CF/verifier.rs:53-236)
  verify_proof                 the harness's verify_proof (BT/harness/src/main.rs:360-391):
                               the commitment root recomputed from the seeded inputs, then
                               the verifier

Every function raises Reject on a failed check.  The oracle computes each
field product, hash and fold; nothing here calls the device library.
"""
import os
import re

import numpy as np

from oracle_ffi import F128_DT, OracleFs, f128_to_ints, ints_to_f128
from ligerito_verifier import Quad, ceil_log2, verify_level_opens

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
MAX_PROOF_BYTES = 500_000
K_LOG, K_SKIP, CONST_POS = 14, 6, 512
DOMAIN = b"flock-bench-v0"


class Reject(Exception):
    pass


def need(cond, what):
    if not cond:
        raise Reject(what)


# ---------------------------------------------------------------------------
# Ligerito ladders: the generated table of the package (one row per (m, profile)).
# ---------------------------------------------------------------------------
def ladders():
    text = open(os.path.join(ROOT, "flock_challenge_multi_amd", "csrc", "ligerito_ladders.inc")).read()
    out = {}
    for line in text.splitlines():
        line = line.strip()
        if not line.startswith("{"):
            continue
        nums = [int(x) for x in re.findall(r"\d+", line)]
        m, prof, log_n, k0, yr, nl = nums[:6]
        lv = [nums[6 + 8 * i: 14 + 8 * i] for i in range(nl)]
        out[(m, prof)] = {
            "log_n": log_n, "initial_k": k0, "yr_log_n": yr,
            "log_inv_rates": [l[0] for l in lv], "recursive_ks": [l[3] for l in lv[1:]],
            "queries": [l[4] for l in lv], "grinding_bits": [l[5] for l in lv],
            "fold_grinding_bits": [l[6] for l in lv], "ood_samples": [l[7] for l in lv],
        }
    return out


# ---------------------------------------------------------------------------
# Deserialisation
# ---------------------------------------------------------------------------
class _Reader:
    def __init__(self, b):
        self.b, self.o = b, 0

    def take(self, n):
        need(self.o + n <= len(self.b), "truncated payload")
        v = self.b[self.o:self.o + n]
        self.o += n
        return v

    def u64(self):
        return int.from_bytes(self.take(8), "little")

    def u32(self):
        return int.from_bytes(self.take(4), "little")

    def f128(self):
        return int.from_bytes(self.take(16), "little")

    def count(self, elem_bytes):
        n = self.u64()
        need(n * elem_bytes <= len(self.b) - self.o, "length prefix past the end")
        return n

    def f128_vec(self):
        return [self.f128() for _ in range(self.count(16))]

    def pairs(self):
        return [(self.f128(), self.f128()) for _ in range(self.count(32))]

    def hashes(self):
        return [bytes(self.take(32)) for _ in range(self.count(32))]

    def u64_vec(self):
        return [self.u64() for _ in range(self.count(8))]

    def rows(self):
        rows = []
        for _ in range(self.count(8)):
            n = self.count(16)
            rows.append(np.frombuffer(bytes(self.take(16 * n)), dtype=F128_DT).copy())
        return rows


def parse_bundle(b):
    need(0 < len(b) <= MAX_PROOF_BYTES, f"proof size {len(b)} outside (0, {MAX_PROOF_BYTES}]")
    need(len(b) >= 7 and b[:5] == b"FLOCK" and b[5] == 4 and b[6] == 2, "bad FLOCK header")
    r = _Reader(b[7:])
    out = {"root": bytes(r.take(32)), "m": r.u64(), "log_inv_rate": r.u64(), "log_batch_size": r.u64()}
    out["profile"] = r.u32()
    out["merkle_hash"] = r.u32()
    need(out["profile"] < 3 and out["merkle_hash"] < 2, "bad enum variant")
    out["zc"] = {"round1_ab": r.f128_vec(), "round1_c": r.f128_vec(), "rounds": r.pairs(),
                 "final_a": r.f128(), "final_b": r.f128(), "final_c": r.f128()}
    out["lc"] = {"rounds": r.pairs(), "z_partial": r.f128_vec()}
    out["s_hat_v"] = [r.f128_vec() for _ in range(r.count(8))]
    lig = {"initial_root": bytes(r.take(32))}
    lig["initial_proof"] = (r.rows(), r.hashes())
    lig["recursive_roots"] = r.hashes()
    lig["recursive_proofs"] = [(r.rows(), r.hashes()) for _ in range(r.count(16))]
    lig["yr"] = r.f128_vec()
    lig["final_proof"] = (r.rows(), r.hashes())
    lig["sumcheck_transcript"] = r.pairs()
    lig["grinding_nonces"] = r.u64_vec()
    lig["ood_values"] = r.f128_vec()
    lig["fold_grinding_nonces"] = r.u64_vec()
    out["lig"] = lig
    need(r.o == len(r.b), "trailing bytes")
    return out


def serialize_bundle(b):
    """The inverse of parse_bundle: bincode fixint of R1csProofBundleLigerito with the
    FLOCK header (written field by field from the struct order of CF/proof.rs:17-22,
    CF/zerocheck.rs:311-324, CF/lincheck.rs:2002-2015, CF/pcs.rs:42-47,
    CF/pcs/ligerito.rs:1021-1048)."""
    out = bytearray(b"FLOCK\x04\x02")
    u64 = lambda v: out.extend(int(v).to_bytes(8, "little"))  # noqa: E731
    f = lambda v: out.extend(int(v).to_bytes(16, "little"))  # noqa: E731

    def fvec(v):
        u64(len(v))
        for x in v:
            f(x)

    def pairs(v):
        u64(len(v))
        for x, y in v:
            f(x)
            f(y)

    def hashes(v):
        u64(len(v))
        for h in v:
            out.extend(h)

    def rows(v):
        u64(len(v))
        for r in v:
            u64(len(r))
            out.extend(np.ascontiguousarray(r, dtype=F128_DT).tobytes())

    out.extend(b["root"])
    u64(b["m"])
    u64(b["log_inv_rate"])
    u64(b["log_batch_size"])
    out.extend(int(b["profile"]).to_bytes(4, "little"))
    out.extend(int(b["merkle_hash"]).to_bytes(4, "little"))
    zc, lc, lig = b["zc"], b["lc"], b["lig"]
    fvec(zc["round1_ab"])
    fvec(zc["round1_c"])
    pairs(zc["rounds"])
    for k in ("final_a", "final_b", "final_c"):
        f(zc[k])
    pairs(lc["rounds"])
    fvec(lc["z_partial"])
    u64(len(b["s_hat_v"]))
    for v in b["s_hat_v"]:
        fvec(v)
    out.extend(lig["initial_root"])
    rows(lig["initial_proof"][0])
    hashes(lig["initial_proof"][1])
    hashes(lig["recursive_roots"])
    u64(len(lig["recursive_proofs"]))
    for rw, hs in lig["recursive_proofs"]:
        rows(rw)
        hashes(hs)
    fvec(lig["yr"])
    rows(lig["final_proof"][0])
    hashes(lig["final_proof"][1])
    pairs(lig["sumcheck_transcript"])
    for k in ("grinding_nonces",):
        u64(len(lig[k]))
        for x in lig[k]:
            u64(x)
    fvec(lig["ood_values"])
    u64(len(lig["fold_grinding_nonces"]))
    for x in lig["fold_grinding_nonces"]:
        u64(x)
    return bytes(out)


# ---------------------------------------------------------------------------
# Field helpers
# ---------------------------------------------------------------------------
def inner(orc, a, b):
    acc = 0
    for x, y in zip(a, b):
        acc ^= orc.mul(x, y)
    return acc


def lagrange_on_nodes(orc, nodes, z):
    """L_i(z) = prod_{j != i} (z + s_j) / (s_i + s_j) (CF/zerocheck/multilinear.rs:336-395)."""
    out = []
    for i, si in enumerate(nodes):
        num = den = 1
        for j, sj in enumerate(nodes):
            if j != i:
                num = orc.mul(num, z ^ sj)
                den = orc.mul(den, si ^ sj)
        out.append(orc.mul(num, orc.inv(den)))
    return out


def phi8_nodes(orc):
    return f128_to_ints(orc.phi8_table())


# ---------------------------------------------------------------------------
# R1CS: symbolic matrices, statement digest, comb vector
# ---------------------------------------------------------------------------
_MATS = {}


def matrices():
    """(a_ptr, a_cols, b_ptr, b_cols) of A_0, B_0 in CSR with ascending columns."""
    if "csr" not in _MATS:
        import test_oracle_witness as tw
        A, B = tw.build_matrices()

        def csr(rows):
            ptr, cols = [0], []
            for v in rows:
                bits = np.unpackbits(np.frombuffer(v.to_bytes((1 << K_LOG) // 8, "little"), np.uint8),
                                     bitorder="little")
                nz = np.nonzero(bits)[0]
                cols.append(nz.astype(np.uint32))
                ptr.append(ptr[-1] + len(nz))
            return np.array(ptr, np.uint32), np.concatenate(cols)

        _MATS["csr"] = csr(A) + csr(B)
    return _MATS["csr"]


def statement_digest(orc, m):
    """BLAKE3 of "flock-r1cs-stmt-v1" || m, k_log, k_skip (u64le) || 0 (RowMajor) ||
    absorb_matrix(A_0) || absorb_matrix(B_0) || absorb_matrix(I)."""
    a_ptr, a_cols, b_ptr, b_cols = matrices()
    k = 1 << K_LOG
    parts = [b"flock-r1cs-stmt-v1", np.array([m, K_LOG, K_SKIP], "<u8").tobytes(), b"\x00"]
    for ptr, cols in ((a_ptr, a_cols), (b_ptr, b_cols)):
        lens = np.diff(ptr.astype(np.int64))
        seq = np.empty(len(cols) + k, np.uint64)
        starts = ptr[:-1].astype(np.int64) + np.arange(k)
        seq[starts] = lens
        mask = np.ones(len(seq), bool)
        mask[starts] = False
        seq[mask] = cols
        parts += [np.array([k, k], "<u8").tobytes(), seq.astype("<u8").tobytes()]
    ident = np.empty(2 * k, np.uint64)
    ident[0::2] = 1
    ident[1::2] = np.arange(k)
    parts += [np.array([k, k], "<u8").tobytes(), ident.astype("<u8").tobytes()]
    return orc.blake3(b"".join(parts))


def quirky_eq(orc, z_skip, x_inner_rest):
    """build_quirky_eq_table (CF/lincheck.rs:4653-4665): index i_skip + 64 i_rest."""
    lam = orc.zc_lagrange_weights(K_SKIP, z_skip)
    er = orc.build_eq(x_inner_rest)
    return orc.mul_many(np.tile(lam, len(er)), np.repeat(er, len(lam)))


def lincheck_comb(orc, alpha, eq_inner):
    """fold_alpha_batched: alpha * A^T eq + B^T eq (CF/lincheck.rs:4683-4720)."""
    a_ptr, a_cols, b_ptr, b_cols = matrices()
    k = 1 << K_LOG
    fa = orc.sparse_fold(a_ptr, a_cols, eq_inner, k)
    fb = orc.sparse_fold(b_ptr, b_cols, eq_inner, k)
    al = ints_to_f128([alpha] * k)
    comb = orc.mul_many(fa, al)
    comb["lo"] ^= fb["lo"]
    comb["hi"] ^= fb["hi"]
    return comb


# ---------------------------------------------------------------------------
# Zerocheck (CF/zerocheck.rs:1642-1800)
# ---------------------------------------------------------------------------
def fixed_challenges(orc):
    phi = phi8_nodes(orc)
    small = [phi[v] for v in (0xF7, 0x53, 0xB5)]
    medium = []
    for i in range(4):
        g = 1 << (1 << i)
        medium.append(orc.mul(g, orc.inv(1 ^ g)))
    return small + medium


def zerocheck_verify(orc, m, zc, ch):
    n_mlv = m - K_SKIP
    need(m >= K_SKIP + 7, "log_n too small")
    need(len(zc["round1_ab"]) == 64 and len(zc["round1_c"]) == 64, "round-1 length")
    need(len(zc["rounds"]) == n_mlv, "multilinear rounds length")
    ch.observe_label(b"flock-zerocheck-v0")
    r_skip = ch.sample_f128_vec(K_SKIP)
    r_outer = ch.sample_f128_vec(m - K_SKIP - 7)
    r = r_skip + fixed_challenges(orc) + r_outer
    ch.observe_f128_slice(zc["round1_ab"])
    ch.observe_f128_slice(zc["round1_c"])
    z = ch.sample_f128()
    phi = phi8_nodes(orc)
    lam = lagrange_on_nodes(orc, phi[64:128], z)
    c_eval = inner(orc, lam, zc["round1_c"])
    need(c_eval == zc["final_c"], "c_eval mismatch")
    # P^AB(z) from the combined polynomial, zero on S (interpolate_at_z_combined).
    combined = [x ^ y for x, y in zip(zc["round1_ab"], zc["round1_c"])]
    comb_z = 0
    for i in range(64):
        si = phi[64 + i]
        num = den = 1
        for j in range(128):
            if j != 64 + i:
                num = orc.mul(num, z ^ phi[j])
                den = orc.mul(den, si ^ phi[j])
        comb_z ^= orc.mul(orc.mul(num, orc.inv(den)), combined[i])
    running = comb_z ^ c_eval
    rhos = []
    for i, (g1, ginf) in enumerate(zc["rounds"]):
        r_eq = r[K_SKIP + i]
        g0 = orc.mul(running ^ orc.mul(r_eq, g1), orc.inv(1 ^ r_eq))
        ch.observe_f128(g1)
        ch.observe_f128(ginf)
        rho = ch.sample_f128()
        rhos.append(rho)
        opr = 1 ^ rho
        running = orc.mul(g0, opr) ^ orc.mul(g1, rho) ^ orc.mul(ginf, orc.mul(rho, opr))
    need(running == orc.mul(zc["final_a"], zc["final_b"]), "zerocheck final check")
    ch.observe_f128(zc["final_a"])
    ch.observe_f128(zc["final_b"])
    return {"z": z, "mlv": rhos, "r_rest": r[K_SKIP:], "a_eval": zc["final_a"], "b_eval": zc["final_b"],
            "c_eval": zc["final_c"]}


# ---------------------------------------------------------------------------
# Lincheck (CF/lincheck.rs:5880-6040)
# ---------------------------------------------------------------------------
def lincheck_verify(orc, m, x_ab, v_a, v_b, lc, ch):
    rest = K_LOG - K_SKIP
    need(len(lc["rounds"]) == rest, "lincheck rounds length")
    need(len(lc["z_partial"]) == 1 << K_SKIP, "z_partial length")
    ch.observe_label(b"flock-lincheck-v0")
    alpha = ch.sample_f128()
    comb = lincheck_comb(orc, alpha, quirky_eq(orc, x_ab["z_skip"], x_ab["x_inner_rest"]))
    target = orc.mul(alpha, v_a) ^ v_b
    beta = ch.sample_f128()
    cp = f128_to_ints(comb[CONST_POS:CONST_POS + 1])[0] ^ beta
    comb[CONST_POS] = (cp & (2**64 - 1), cp >> 64)
    target ^= beta
    running = target
    rr = []
    for e1, einf in lc["rounds"]:
        ch.observe_f128(e1)
        ch.observe_f128(einf)
        ri = ch.sample_f128()
        e0 = running ^ e1
        c1 = e0 ^ e1 ^ einf
        running = orc.mul(einf, orc.mul(ri, ri)) ^ orc.mul(c1, ri) ^ e0
        h = len(comb) // 2
        lo, hi = comb[:h].copy(), comb[h:].copy()
        d = lo.copy()
        d["lo"] ^= hi["lo"]
        d["hi"] ^= hi["hi"]
        t = orc.mul_many(d, ints_to_f128([ri] * h))
        lo["lo"] ^= t["lo"]
        lo["hi"] ^= t["hi"]
        comb = lo
        rr.append(ri)
    ch.observe_f128_slice(lc["z_partial"])
    need(inner(orc, f128_to_ints(comb), lc["z_partial"]) == running, "lincheck final check")
    r_skip = ch.sample_f128()
    w = inner(orc, f128_to_ints(orc.zc_lagrange_weights(K_SKIP, r_skip)), lc["z_partial"])
    return {"r_inner_skip": r_skip, "r_inner_rest": rr[::-1], "w": w}


# ---------------------------------------------------------------------------
# Ring switch and the DP24 succinct evaluation (CF/pcs/ring_switch.rs, tensor_algebra.rs)
# ---------------------------------------------------------------------------
def ta_transpose(e):
    raw = b"".join(int(v).to_bytes(16, "little") for v in e)
    bits = np.unpackbits(np.frombuffer(raw, np.uint8), bitorder="little").reshape(128, 128)
    t = np.packbits(np.ascontiguousarray(bits.T), axis=1, bitorder="little")
    return [int.from_bytes(t[i].tobytes(), "little") for i in range(128)]


def ta_scale(orc, e, s):
    return f128_to_ints(orc.mul_many(ints_to_f128(e), ints_to_f128([s] * 128)))


def eval_rs_eq_prefix(orc, z_vals, query_prefix):
    ev = [1] + [0] * 127
    for zi, qi in zip(z_vals, query_prefix):
        vert = ta_scale(orc, ev, zi)
        hztl = ta_transpose(ta_scale(orc, ta_transpose(ev), qi))
        ev = [a ^ b ^ c for a, b, c in zip(ev, vert, hztl)]
    return ev


def eval_rs_eq_finish_binary(orc, prefix, z_suffix, y_bits, eq_r):
    s = 1
    for j, zi in enumerate(z_suffix):
        s = orc.mul(s, zi if (y_bits >> j) & 1 else 1 ^ zi)
    return inner(orc, eq_r, ta_transpose(ta_scale(orc, prefix, s)))


def ring_switch_verify_succinct(orc, claim, z_skip, x_outer, s_hat_v, ch):
    need(len(s_hat_v) == 128, "s_hat_v length")
    ch.observe_label(b"flock-ring-switch-v0")
    ch.observe_f128_slice(s_hat_v)
    need(orc.rs_claim_check(z_skip, x_outer[0], s_hat_v) == claim, "ring-switch claim mismatch")
    r_dp = ch.sample_f128_vec(7)
    eq_r = f128_to_ints(orc.build_eq(r_dp))
    return {"sumcheck_claim": inner(orc, orc.rs_transpose(s_hat_v), eq_r), "eq_r": eq_r}


# ---------------------------------------------------------------------------
# Ligerito, succinct (CF/pcs/ligerito.rs:11321-11834)
# ---------------------------------------------------------------------------
def hash_array(hashes):
    return np.frombuffer(b"".join(hashes), np.uint8).reshape(-1, 32)


def enforced_sum(orc, rows, v_challenges, queries, alpha):
    eq = orc.build_eq(v_challenges)
    aw = f128_to_ints(orc.build_eq(alpha))[:len(queries)]
    s = 0
    for i, row in enumerate(rows):
        need(len(row) == len(eq), "opened row length")
        p = orc.mul_many(row, eq)
        s ^= orc.mul(aw[i], int(np.bitwise_xor.reduce(p["lo"])) | (int(np.bitwise_xor.reduce(p["hi"])) << 64))
    return s


def next_s(orc, s, s_at_root):
    return orc.mul(s, s) ^ orc.mul(s_at_root, s)


def induce_residual(orc, log_msg_cols, queries, alpha, ris_for_basis, yr_log_n):
    need(len(ris_for_basis) + yr_log_n == log_msg_cols, "residual split")
    sks = f128_to_ints(orc.eval_sk_at_vks(log_msg_cols))
    inv = [orc.inv(v) if v else 0 for v in sks]
    aw = f128_to_ints(orc.build_eq(alpha))[:len(queries)]
    per = []
    for q in queries:
        s = [q]
        for k in range(1, log_msg_cols):
            s.append(next_s(orc, s[k - 1], sks[k - 1]))
        s = [orc.mul(s[k], inv[k]) for k in range(log_msg_cols)]
        pre = 1
        for k in range(len(ris_for_basis)):
            pre = orc.mul(pre, 1 ^ orc.mul(ris_for_basis[k], 1 ^ s[k]))
        per.append((pre, s[len(ris_for_basis):]))
    out = []
    for y in range(1 << yr_log_n):
        tot = 0
        for i, (pre, w) in enumerate(per):
            suf = 1
            for j in range(yr_log_n):
                if (y >> j) & 1:
                    suf = orc.mul(suf, w[j])
            tot ^= orc.mul(aw[i], orc.mul(pre, suf))
        out.append(tot)
    return out


def ligerito_verify_succinct(orc, cfg, lig, log_n, target, expected_root, eval_b_residual, ch):
    k0, ks, rates = cfg["initial_k"], cfg["recursive_ks"], cfg["log_inv_rates"]
    r = len(ks)
    need(r >= 1 and len(rates) == r + 1, "config shape")
    need(lig["initial_root"] == expected_root, "initial root")
    ch.observe_label(b"flock-ligerito-basis-v0")
    ch.observe_f128(target)
    ch.observe_bytes(lig["initial_root"])
    tx = list(lig["sumcheck_transcript"])
    fold_nonces = list(lig["fold_grinding_nonces"])
    oods = list(lig["ood_values"])
    nonces = lig["grinding_nonces"]

    def next_msg():
        need(tx, "sumcheck transcript too short")
        msg = tx.pop(0)
        ch.observe_f128(msg[0])
        ch.observe_f128(msg[1])
        return msg

    def fold_bits(level):
        return cfg["fold_grinding_bits"][level] if level < len(cfg["fold_grinding_bits"]) else 0

    def ood_count(level):
        return cfg["ood_samples"][level] if level < len(cfg["ood_samples"]) else 0

    t_r = target
    quad = Quad.from_msg(orc, next_msg(), t_r)
    need(ood_count(0) == 0, "L0 takes no OOD samples")
    ood_ctxs, ris = [], []

    def fold_rounds(k, level):
        nonlocal quad, t_r
        rs = []
        for j in range(k):
            bits = max(fold_bits(level) - j, 0)
            if bits > 0:
                need(fold_nonces, "fold grinding nonces missing")
                need(ch.verify_pow(fold_nonces.pop(0), bits), "fold grind")
            ri = ch.sample_f128()
            rs.append(ri)
            ris.append(ri)
            t_r = quad.eval(ri)
            quad = Quad.from_msg(orc, next_msg(), t_r)
        return rs

    def ood(level, n_vars):
        nonlocal quad, t_r
        for _ in range(ood_count(level)):
            z = ch.sample_f128_vec(n_vars)
            need(oods, "OOD values missing")
            y = oods.pop(0)
            ch.observe_f128(y)
            intro = Quad.from_msg(orc, next_msg(), y)
            beta = ch.sample_f128()
            quad = quad.fold(intro, beta)
            t_r ^= orc.mul(beta, y)
            ood_ctxs.append((z, len(ris), beta))

    def introduce(s):
        nonlocal quad, t_r
        intro = Quad.from_msg(orc, next_msg(), s)
        beta = ch.sample_f128()
        quad = quad.fold(intro, beta)
        t_r ^= orc.mul(beta, s)
        return beta

    r_lane = fold_rounds(k0, 0)
    need(lig["recursive_roots"], "no recursive roots")
    prev_root = lig["recursive_roots"][0]
    ch.observe_bytes(prev_root)
    n_cur = log_n - k0
    ood(1, n_cur)
    need(len(nonces) > 0 and ch.verify_pow(nonces[0], cfg["grinding_bits"][0]), "query grind 0")
    nonce_idx = 1
    block0 = 1 << (n_cur + rates[0])
    q0 = ch.sample_distinct_queries(block0, cfg["queries"][0])
    a0 = ch.sample_f128_vec(ceil_log2(cfg["queries"][0]))
    rows0, mp0 = lig["initial_proof"]
    need(len(rows0) == len(q0) and all(len(x) == 1 << k0 for x in rows0), "L0 opened rows shape")
    need(verify_level_opens(orc, lig["initial_root"], block0, q0, np.array(rows0), hash_array(mp0)), "L0 openings")
    es0 = enforced_sum(orc, rows0, r_lane, q0, a0)
    beta0 = introduce(es0)
    levels = [(n_cur, q0, a0, k0, beta0)]
    prev_k, prev_rate = ks[0], rates[1]
    prev_cols = n_cur - prev_k
    next_root, rp_idx = 1, 0
    for i in range(r):
        need(n_cur >= ks[i], "ladder folds past the variables")
        level_rs = fold_rounds(ks[i], i + 1)
        n_cur -= ks[i]
        prev_block = 1 << (prev_cols + prev_rate)
        if i == r - 1:
            need(not tx and not oods and not fold_nonces, "unconsumed proof values")
            yr = lig["yr"]
            need(len(yr) == 1 << n_cur, "yr length")
            for v in yr:
                ch.observe_f128(v)
            need(nonce_idx < len(nonces) and ch.verify_pow(nonces[nonce_idx], cfg["grinding_bits"][i + 1]),
                 "final query grind")
            ql = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
            al = ch.sample_f128_vec(ceil_log2(cfg["queries"][i + 1]))
            rows, mp = lig["final_proof"]
            need(len(rows) == len(ql) and all(len(x) == 1 << prev_k for x in rows), "final opened rows shape")
            need(verify_level_opens(orc, prev_root, prev_block, ql, np.array(rows), hash_array(mp)), "final openings")
            es = enforced_sum(orc, rows, level_rs, ql, al)
            beta_last = ch.sample_f128()
            t_r ^= orc.mul(beta_last, es)
            levels.append((n_cur, ql, al, len(ris), beta_last))
            yr_log_n = n_cur
            combined = eval_b_residual(ris, yr_log_n)
            need(len(combined) == len(yr), "residual length")
            for cols, qs, al_, start, beta in levels:
                res = induce_residual(orc, cols, qs, al_, ris[start:start + cols - yr_log_n], yr_log_n)
                combined = [c ^ orc.mul(beta, x) for c, x in zip(combined, res)]
            for z, start, beta in ood_ctxs:
                need(len(z) >= yr_log_n and start + len(z) - yr_log_n <= len(ris), "OOD residual shape")
                folded = len(z) - yr_log_n
                sc = beta
                for b in range(folded):
                    sc = orc.mul(sc, 1 ^ z[b] ^ ris[start + b])
                tail = f128_to_ints(orc.build_eq(z[folded:]))
                combined = [c ^ orc.mul(sc, t) for c, t in zip(combined, tail)]
            return inner(orc, yr, combined) == t_r
        need(next_root < len(lig["recursive_roots"]), "recursive roots missing")
        root_next = lig["recursive_roots"][next_root]
        next_root += 1
        ch.observe_bytes(root_next)
        ood(i + 2, n_cur)
        need(nonce_idx < len(nonces) and ch.verify_pow(nonces[nonce_idx], cfg["grinding_bits"][i + 1]),
             "query grind")
        nonce_idx += 1
        qi = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
        ai = ch.sample_f128_vec(ceil_log2(cfg["queries"][i + 1]))
        need(rp_idx < len(lig["recursive_proofs"]), "recursive proofs missing")
        rows, mp = lig["recursive_proofs"][rp_idx]
        rp_idx += 1
        need(len(rows) == len(qi) and all(len(x) == 1 << prev_k for x in rows), "opened rows shape")
        need(verify_level_opens(orc, prev_root, prev_block, qi, np.array(rows), hash_array(mp)), "level openings")
        es = enforced_sum(orc, rows, level_rs, qi, ai)
        beta = introduce(es)
        levels.append((n_cur, qi, ai, len(ris), beta))
        prev_root = root_next
        need(n_cur >= ks[i + 1] if i + 1 < r else True, "ladder")
        prev_k, prev_rate = ks[i + 1], rates[i + 2]
        prev_cols = n_cur - prev_k
    raise Reject("unreachable")


# ---------------------------------------------------------------------------
# PCS and the R1CS proof
# ---------------------------------------------------------------------------
def verify_pcs(orc, root, m, claims, s_hat_vs, lig, cfg, ch):
    """claims: [(value, z_skip, x_full)], x_full = x_inner_rest ++ x_outer."""
    need(len(s_hat_vs) == len(claims), "ring-switch count")
    ch.observe_label(b"flock-pcs-open-batch-v0")
    outs = [ring_switch_verify_succinct(orc, v, zs, x, s, ch) for (v, zs, x), s in zip(claims, s_hat_vs)]
    gammas = [ch.sample_f128() for _ in claims]
    target = 0
    for o, g in zip(outs, gammas):
        target ^= orc.mul(g, o["sumcheck_claim"])
    log_n = m - 7

    def eval_b_residual(ris, yr_log_n):
        n = len(ris)
        prefixes = [eval_rs_eq_prefix(orc, x[1:1 + n], ris) for (_, _, x) in claims]
        out = []
        for y in range(1 << yr_log_n):
            s = 0
            for (_, _, x), o, g, pre in zip(claims, outs, gammas, prefixes):
                s ^= orc.mul(g, eval_rs_eq_finish_binary(orc, pre, x[1 + n:], y, o["eq_r"]))
            out.append(s)
        return out

    need(ligerito_verify_succinct(orc, cfg, lig, log_n, target, root, eval_b_residual, ch), "Ligerito")


def verify_r1cs(orc, bundle, domain=DOMAIN):
    """verify_ligerito: bind_statement, zerocheck, lincheck, the batched PCS opening."""
    m = bundle["m"]
    cfg = ladders().get((m, bundle["profile"]))
    need(cfg is not None and bundle["log_batch_size"] == 6 and bundle["log_inv_rate"] == cfg["log_inv_rates"][0],
         "PCS parameters")
    ch = OracleFs(orc, domain)
    ch.observe_label(b"flock-r1cs-v0")
    ch.observe_bytes(statement_digest(orc, m))
    ch.observe_bytes(bundle["root"])
    zc = zerocheck_verify(orc, m, bundle["zc"], ch)
    rest = K_LOG - K_SKIP
    x_ab = {"z_skip": zc["z"], "x_inner_rest": zc["mlv"][:rest], "x_outer": zc["mlv"][rest:]}
    lc = lincheck_verify(orc, m, x_ab, zc["a_eval"], zc["b_eval"], bundle["lc"], ch)
    claims = [(lc["w"], lc["r_inner_skip"], lc["r_inner_rest"] + x_ab["x_outer"]),
              (zc["c_eval"], zc["z"], zc["r_rest"])]
    verify_pcs(orc, bundle["root"], m, claims, bundle["s_hat_v"], bundle["lig"], cfg, ch)
    return zc, lc


def expected_root(orc, comps, n_blocks_log, threads=8):
    """pcs::commit of the witness the harness regenerates (BT/harness/src/main.rs:365-374)."""
    z, _, _, _ = orc.witness_mt(comps, n_blocks_log, threads, stripe=False)
    msg = z.view(F128_DT)
    _, _, root = orc.commit_mt(msg, 64, 1, threads)
    return root


def verify_proof(orc, proof_bytes, comps, n_blocks_log, domain=DOMAIN):
    """The harness's check: header and size, the commitment's root and parameters against
    the regenerated witness, then the verifier.  Raises Reject."""
    b = parse_bundle(proof_bytes)
    need(b["m"] == 14 + n_blocks_log and b["merkle_hash"] == 1 and b["profile"] == 0, "commitment parameters")
    need(b["root"] == expected_root(orc, comps, n_blocks_log), "commitment root differs from the witness")
    return verify_r1cs(orc, b, domain)
