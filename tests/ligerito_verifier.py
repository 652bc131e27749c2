"""Test infrastructure: recursive_verifier_with_basis (CF/pcs/ligerito.rs:11836-12228,
CF = crates/flock-core/src/) restated over the oracle's primitives (transcript
byte log, partial evaluation, basis induction, BLAKE3 leaves and Merkle
multi-proof verification, PoW check).  It checks a proof produced by the
device prover: the transcript replay, every fold grind and query grind, every
opened row against its level's root, the running sumcheck claim through
RoundQuad (:4929-4956), and at the last level <yr, combined basis> = T with
the combined basis rebuilt from the initial basis, every induced basis and
every OOD eq table."""
import numpy as np

from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128


def ceil_log2(n):
    return 0 if n <= 1 else (n - 1).bit_length()


class Quad:
    """RoundQuad: c + b X + a X^2 with c = u_0, b = T + u_2, a = u_2."""

    def __init__(self, orc, c, b, a):
        self.orc, self.c, self.b, self.a = orc, c, b, a

    @classmethod
    def from_msg(cls, orc, msg, t):
        return cls(orc, msg[0], t ^ msg[1], msg[1])

    def eval(self, r):
        m = self.orc.mul
        return self.c ^ m(r, self.b) ^ m(m(r, r), self.a)

    def fold(self, other, alpha):
        m = self.orc.mul
        return Quad(self.orc, self.c ^ m(alpha, other.c), self.b ^ m(alpha, other.b), self.a ^ m(alpha, other.a))


def verify_level_opens(orc, root, block_len, queries, rows, proof):
    """verify_level_opens (:12441-12480): leaf = BLAKE3 chunk CV of the row bytes."""
    if len(rows) != len(queries):
        return False
    leaves = np.frombuffer(b"".join(orc.leaf_cv(np.ascontiguousarray(row).tobytes()) for row in rows),
                           np.uint8).reshape(-1, 32)
    return orc.multi_verify(root, block_len, list(queries), leaves, np.asarray(proof, np.uint8).reshape(-1, 32))


def verify(orc, cfg, proof, b_initial, target, expected_root, ch):
    """proof: a dict with the LigeritoProof fields in host form (rows as numpy
    (n_queries, row_len) F128 arrays, Merkle proofs as (k, 32) uint8 arrays)."""
    log_n = int(len(b_initial)).bit_length() - 1
    k0, r = cfg["initial_k"], len(cfg["recursive_ks"])
    if proof["initial_root"] != expected_root:
        return False
    ch.observe_label(b"flock-ligerito-basis-v0")
    ch.observe_f128(target)
    ch.observe_bytes(proof["initial_root"])
    tx = list(proof["sumcheck_transcript"])
    fold_nonces = list(proof["fold_grinding_nonces"])
    nonces = list(proof["grinding_nonces"])
    oods = list(proof["ood_values"])
    t_r = target

    def next_msg():
        if not tx:
            raise IndexError
        msg = tx.pop(0)
        ch.observe_f128(msg[0])
        ch.observe_f128(msg[1])
        return msg

    try:
        quad = Quad.from_msg(orc, next_msg(), t_r)
        ris = []

        def fold_rounds(k, level):
            nonlocal quad, t_r
            rs = []
            for j in range(k):
                bits = max(cfg["fold_grinding_bits"][level] - j, 0)
                if bits > 0 and not ch.verify_pow(fold_nonces.pop(0), bits):
                    raise ValueError("fold grind")
                ri = ch.sample_f128()
                rs.append(ri)
                ris.append(ri)
                t_r = quad.eval(ri)
                quad = Quad.from_msg(orc, next_msg(), t_r)
            return rs

        ood_bases = []

        def ood(level, n_vars):
            nonlocal quad, t_r
            for _ in range(cfg["ood_samples"][level]):
                z = ch.sample_f128_vec(n_vars)
                y = oods.pop(0)
                ch.observe_f128(y)
                intro = Quad.from_msg(orc, next_msg(), y)
                beta = ch.sample_f128()
                quad = quad.fold(intro, beta)
                t_r ^= orc.mul(beta, y)
                ood_bases.append((orc.build_eq(z), len(ris), beta))

        fold_rounds(k0, 0)
        roots = list(proof["recursive_roots"])
        prev_root = roots.pop(0)
        ch.observe_bytes(prev_root)
        n_cur = log_n - k0
        ood(1, n_cur)
        if not ch.verify_pow(nonces.pop(0), cfg["grinding_bits"][0]):
            return False
        block0 = 1 << (n_cur + cfg["log_inv_rates"][0])
        q0 = ch.sample_distinct_queries(block0, cfg["queries"][0])
        a0 = ch.sample_f128_vec(ceil_log2(cfg["queries"][0]))
        rows0, mp0 = proof["initial_proof"]
        if not verify_level_opens(orc, proof["initial_root"], block0, q0, rows0, mp0):
            return False
        basis0, sum0 = orc.induce(n_cur, rows0, ris[:k0], q0, a0)
        intro = Quad.from_msg(orc, next_msg(), sum0)
        beta = ch.sample_f128()
        quad = quad.fold(intro, beta)
        t_r ^= orc.mul(beta, sum0)
        bases = [(b_initial, 0, 1), (basis0, k0, beta)]
        prev_k, prev_rate = cfg["recursive_ks"][0], cfg["log_inv_rates"][1]
        prev_cols = n_cur - prev_k
        rproofs = list(proof["recursive_proofs"])
        for i in range(r):
            level_rs = fold_rounds(cfg["recursive_ks"][i], i + 1)
            n_cur -= cfg["recursive_ks"][i]
            prev_block = 1 << (prev_cols + prev_rate)
            if i == r - 1:
                if tx or oods or fold_nonces:
                    return False
                yr = proof["yr"]
                if len(yr) != 1 << n_cur:
                    return False
                for v in yr:
                    ch.observe_f128(v)
                if not ch.verify_pow(nonces.pop(0), cfg["grinding_bits"][i + 1]):
                    return False
                ql = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
                al = ch.sample_f128_vec(ceil_log2(cfg["queries"][i + 1]))
                rows, mp = proof["final_rows"], proof["final_merkle"]
                if not verify_level_opens(orc, prev_root, prev_block, ql, rows, mp):
                    return False
                basis, s = orc.induce(n_cur, rows, level_rs, ql, al)
                beta = ch.sample_f128()
                t_r ^= orc.mul(beta, s)
                bases.append((basis, len(ris), beta))
                combined = [0] * len(yr)
                for bvec, start, sep in bases + ood_bases:
                    res = f128_to_ints(orc.partial_eval_lsb(bvec, ris[start:]))
                    if len(res) != len(yr):
                        return False
                    combined = [c ^ orc.mul(sep, x) for c, x in zip(combined, res)]
                inner = 0
                for y, c in zip(yr, combined):
                    inner ^= orc.mul(y, c)
                return inner == t_r and not nonces
            root_next = roots.pop(0)
            ch.observe_bytes(root_next)
            ood(i + 2, n_cur)
            if not ch.verify_pow(nonces.pop(0), cfg["grinding_bits"][i + 1]):
                return False
            qi = ch.sample_distinct_queries(prev_block, cfg["queries"][i + 1])
            ai = ch.sample_f128_vec(ceil_log2(cfg["queries"][i + 1]))
            rows, mp = rproofs.pop(0)
            if not verify_level_opens(orc, prev_root, prev_block, qi, rows, mp):
                return False
            basis, s = orc.induce(n_cur, rows, level_rs, qi, ai)
            intro = Quad.from_msg(orc, next_msg(), s)
            beta = ch.sample_f128()
            quad = quad.fold(intro, beta)
            t_r ^= orc.mul(beta, s)
            bases.append((basis, len(ris), beta))
            prev_root = root_next
            prev_k, prev_rate = cfg["recursive_ks"][i + 1], cfg["log_inv_rates"][i + 2]
            prev_cols = n_cur - prev_k
    except (IndexError, ValueError):
        return False
    return False
