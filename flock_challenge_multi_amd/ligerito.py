"""The recursive Ligerito prover over a given basis (CF = crates/flock-core/src/):
recursive_prover_with_basis (CF/pcs/ligerito.rs:9996-10026, body :10185-11040),
sequenced on the host over the device entry points.

Every data-parallel step runs on the device (sumcheck rounds and folds, the
recursive ligero_commits, OOD samples, basis induction, row gathers and
Merkle multi-proofs, PoW grinding); the host keeps the Fiat-Shamir transcript
and the proof's small values.  The transcript order is the reference's, as its
verifier replays it (recursive_verifier_with_basis, :11836-12228): label
"flock-ligerito-basis-v0", target, initial root, the first message; per level
the fold rounds (each after its fold grind), the next root, the OOD samples,
the query grind, the queries and alpha, the induced basis's message and beta;
at the last level yr, the grind and the final queries.  Grinds return the
smallest valid nonce (grind_pow; the reference's grind_pow_any_valid may
return any valid one, which its verifier accepts alike).

The L0 commitment is the caller's (the PCS commit of the packed witness:
codeword and tree kept on the device), f is the packed witness and b the
combined basis of the ring switch (flk_pcs_combined_basis_and_target).
"""
from dataclasses import dataclass, field
from typing import List

import torch

from . import hotpath as hp


@dataclass
class ProverConfig:
    """ligerito::ProverConfig's protocol fields (CF/pcs/ligerito.rs:80-110):
    recursive_steps = len(recursive_ks); per level 0..r: log_inv_rates, queries,
    grinding_bits, fold_grinding_bits and ood_samples (ood_samples[0] = 0)."""
    initial_k: int
    recursive_ks: List[int]
    log_inv_rates: List[int]
    queries: List[int]
    grinding_bits: List[int]
    fold_grinding_bits: List[int]
    ood_samples: List[int]

    def check(self, log_n):
        r = len(self.recursive_ks)
        if r < 1:
            raise ValueError("ProverConfig: at least one recursive step")
        for name in ("log_inv_rates", "queries", "grinding_bits", "fold_grinding_bits", "ood_samples"):
            if len(getattr(self, name)) != r + 1:
                raise ValueError(f"ProverConfig: {name} needs {r + 1} entries")
        if self.ood_samples[0]:
            raise ValueError("ProverConfig: L0 takes no OOD samples")
        if log_n - self.initial_k - sum(self.recursive_ks) < 1:
            # every fold needs two elements per output, and yr keeps at least one variable
            raise ValueError("ProverConfig: the folds leave no variable for yr")


@dataclass
class RecursiveProof:
    opened_rows: torch.Tensor      # (n_queries * row_len, 2) int64, query-major
    merkle_proof: torch.Tensor     # (n_hashes, 32) uint8


@dataclass
class FinalProof:
    yr: List[int]
    opened_rows: torch.Tensor
    merkle_proof: torch.Tensor


@dataclass
class LigeritoProof:
    """LigeritoProof (CF/pcs/ligerito.rs:1030-1050)."""
    initial_root: bytes
    initial_proof: RecursiveProof
    recursive_roots: List[bytes] = field(default_factory=list)
    recursive_proofs: List[RecursiveProof] = field(default_factory=list)
    final_proof: FinalProof = None
    sumcheck_transcript: List[tuple] = field(default_factory=list)
    grinding_nonces: List[int] = field(default_factory=list)
    ood_values: List[int] = field(default_factory=list)
    fold_grinding_nonces: List[int] = field(default_factory=list)


def ceil_log2(n):
    """ceil_log2 (CF/pcs/ligerito.rs:1338-1344)."""
    return 0 if n <= 1 else (n - 1).bit_length()


def _ints(t):
    h = t.detach().cpu().numpy().reshape(-1, 2).view("uint64")
    return [int(lo) | (int(hi) << 64) for lo, hi in h]


def _root(tree):
    return bytes(tree[-1].cpu().numpy().tobytes())


def _open(pd, row_len, queries):
    rows = hp.gather_rows(pd.codeword, row_len, queries)
    proof = hp.merkle_multi_proof(pd.merkle_tree, pd.codeword.shape[0] // row_len, queries)
    return RecursiveProof(rows, proof)


def recursive_prover_with_basis(config, f, b, target, l0_codeword, l0_tree, challenger):
    """recursive_prover_with_basis: proves <f, b> = target against the L0
    commitment (l0_codeword: 2^(log_n - initial_k + log_inv_rates[0]) rows of
    2^initial_k elements, l0_tree its Merkle tree).  f and b are device tensors
    of 2^log_n F128; the transcript is an hp.FsChallenger."""
    n = hp._check_f128(f, "f")
    if hp._check_f128(b, "b") != n or not hp._pow2(n):
        raise ValueError("recursive_prover_with_basis: f and b must have equal power-of-two length")
    log_n = hp._log2(n)
    config.check(log_n)
    k0, r = config.initial_k, len(config.recursive_ks)
    block_len_0 = 1 << (log_n - k0 + config.log_inv_rates[0])
    if l0_codeword.shape[0] != block_len_0 << k0:
        raise ValueError("recursive_prover_with_basis: the L0 codeword does not match the config")
    ch = challenger
    proof = LigeritoProof(initial_root=_root(l0_tree), initial_proof=None)

    def observe_msg(msg):
        u0, u2 = _ints(msg) if isinstance(msg, torch.Tensor) else msg
        ch.observe_f128(u0)
        ch.observe_f128(u2)

    def fold_rounds(sp, k, level):
        rs = []
        for j in range(k):
            bits = max(config.fold_grinding_bits[level] - j, 0)
            if bits:
                proof.fold_grinding_nonces.append(ch.grind_pow(bits))
            ri = ch.sample_f128()
            rs.append(ri)
            observe_msg(sp.fold(ri))
        return rs

    def ood(sp, level):
        start = len(sp.transcript)
        ys = sp.ood(ch, config.ood_samples[level])
        proof.ood_values.extend(ys)
        return start

    def introduce(sp, basis, enforced_sum):
        observe_msg(sp.introduce_new(basis, enforced_sum))
        sp.glue(ch.sample_f128())

    ch.observe_label(b"flock-ligerito-basis-v0")
    ch.observe_f128(target)
    ch.observe_bytes(proof.initial_root)
    sp, msg = hp.SumcheckProver.new(f, b, target)
    observe_msg(msg)
    r_lane = fold_rounds(sp, k0, 0)

    # L1 commit of the folded f
    n_cur = log_n - k0
    k_next = config.recursive_ks[0]
    prev = hp.ligero_commit(sp.f, n_cur - k_next, k_next, config.log_inv_rates[1])
    prev_k, prev_block = k_next, 1 << (n_cur - k_next + config.log_inv_rates[1])
    proof.recursive_roots.append(_root(prev.merkle_tree))
    ch.observe_bytes(proof.recursive_roots[-1])
    ood(sp, 1)
    proof.grinding_nonces.append(ch.grind_pow(config.grinding_bits[0]))
    q0 = ch.sample_distinct_queries(block_len_0, config.queries[0])
    alpha0 = ch.sample_f128_vec(ceil_log2(config.queries[0]))
    proof.initial_proof = _open(hp.ProverData(l0_codeword, l0_tree), 1 << k0, q0)
    qd = torch.tensor(q0, dtype=torch.int64, device=f.device)
    basis, esum = hp.induce_sumcheck_poly(n_cur, proof.initial_proof.opened_rows, r_lane, qd, alpha0)
    introduce(sp, basis, _ints(esum)[0])

    for i in range(r):
        level_rs = fold_rounds(sp, config.recursive_ks[i], i + 1)
        n_cur -= config.recursive_ks[i]
        if i == r - 1:
            yr = _ints(sp.f)
            for v in yr:
                ch.observe_f128(v)
            proof.grinding_nonces.append(ch.grind_pow(config.grinding_bits[i + 1]))
            ql = ch.sample_distinct_queries(prev_block, config.queries[i + 1])
            op = _open(prev, 1 << prev_k, ql)
            proof.final_proof = FinalProof(yr, op.opened_rows, op.merkle_proof)
            proof.sumcheck_transcript = [tuple(_ints(m)) if isinstance(m, torch.Tensor) else tuple(m)
                                         for m in sp.transcript]
            return proof
        k_next = config.recursive_ks[i + 1]
        nxt = hp.ligero_commit(sp.f, n_cur - k_next, k_next, config.log_inv_rates[i + 2])
        proof.recursive_roots.append(_root(nxt.merkle_tree))
        ch.observe_bytes(proof.recursive_roots[-1])
        ood(sp, i + 2)
        proof.grinding_nonces.append(ch.grind_pow(config.grinding_bits[i + 1]))
        qi = ch.sample_distinct_queries(prev_block, config.queries[i + 1])
        alpha = ch.sample_f128_vec(ceil_log2(config.queries[i + 1]))
        op = _open(prev, 1 << prev_k, qi)
        proof.recursive_proofs.append(op)
        qd = torch.tensor(qi, dtype=torch.int64, device=f.device)
        basis, esum = hp.induce_sumcheck_poly(n_cur, op.opened_rows, level_rs, qd, alpha)
        introduce(sp, basis, _ints(esum)[0])
        prev, prev_k, prev_block = nxt, k_next, 1 << (n_cur - k_next + config.log_inv_rates[i + 2])
    raise AssertionError("unreachable")
