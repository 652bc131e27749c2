"""ctypes wrapper of the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is test infrastructure: only tests, the smoke check and the
benchmark's CPU-baseline leg load it.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
LIB = os.path.join(ORACLE_DIR, "liboracle.so")

F128_DT = np.dtype([("lo", "<u8"), ("hi", "<u8")])
# One BLAKE3 compression input (112 bytes): cv, message words, counter,
# block_len, flags.
COMPRESSION_DT = np.dtype([("cv", "<u4", 8), ("m", "<u4", 16), ("counter", "<u8"),
                           ("block_len", "<u4"), ("flags", "<u4")])
assert COMPRESSION_DT.itemsize == 112


class COrcF128(ctypes.Structure):
    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def ints_to_f128(vals):
    a = np.zeros(len(vals), dtype=F128_DT)
    for i, v in enumerate(vals):
        a[i]["lo"] = v & (2**64 - 1)
        a[i]["hi"] = v >> 64
    return a


def f128_to_ints(a):
    return [int(x["lo"]) | (int(x["hi"]) << 64) for x in a]


class Oracle:
    def __init__(self):
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
        L = ctypes.CDLL(LIB)
        V, P, SZ, U64 = ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64
        F = COrcF128
        sig = {
            "orc_f128_mul": (F, [F, F]),
            "orc_f128_square": (F, [F]),
            "orc_f128_inv": (F, [F]),
            "orc_f128_mul_by_x": (F, [F]),
            "orc_clmul64": (None, [U64, U64, P, P]),
            "orc_clmul64_bitwise": (None, [U64, U64, P, P]),
            "orc_f128_mul_many": (None, [P, P, P, SZ]),
            "orc_f8_mul": (ctypes.c_uint8, [ctypes.c_uint8, ctypes.c_uint8]),
            "orc_f8_inv": (ctypes.c_uint8, [ctypes.c_uint8]),
            "orc_phi8_table": (None, [P]),
            "orc_fill_f128": (None, [U64, P, SZ]),
            "orc_ntt_twiddles": (ctypes.c_int, [SZ, P]),
            "orc_ntt_forward_interleaved": (ctypes.c_int, [SZ, P, SZ, SZ, SZ]),
            "orc_ntt_inverse": (ctypes.c_int, [SZ, P, SZ]),
            "orc_ntt_rs_encode_interleaved": (ctypes.c_int, [SZ, P, SZ, P, SZ, SZ]),
            "orc_subspace_poly_hat": (F, [SZ, F]),
            "orc_blake3_hash": (None, [P, SZ, P]),
            "orc_blake3_leaf_cv": (None, [P, SZ, P]),
            "orc_blake3_parent_cv": (None, [P, P, P]),
            "orc_blake3_parent_root": (None, [P, P, P]),
            "orc_merkle_tree": (ctypes.c_int, [P, SZ, SZ, P]),
            "orc_merkle_proof": (ctypes.c_int, [P, SZ, SZ, P]),
            "orc_merkle_verify": (ctypes.c_int, [P, P, SZ, P, SZ]),
            "orc_commit": (ctypes.c_int, [P, SZ, SZ, SZ, P, P, P]),
            "orc_commit_mt": (ctypes.c_int, [P, SZ, SZ, SZ, P, P, P, SZ]),
            "orc_build_eq": (None, [P, SZ, P]),
            "orc_partial_eval_lsb": (None, [P, SZ, P, SZ, P]),
            "orc_sc_round_msg": (None, [P, P, SZ, P]),
            "orc_sc_round_msg_and_eval": (None, [P, P, SZ, P]),
            "orc_sc_fold_and_msg": (None, [P, P, SZ, F, P, P, P]),
            "orc_sc_glue": (None, [P, P, SZ, F]),
            "orc_eval_sk_at_vks": (None, [SZ, P]),
            "orc_induce_sumcheck_poly": (ctypes.c_int, [SZ, P, SZ, SZ, P, SZ, P, P, SZ, P, P]),
            "orc_lc_partial_fold": (ctypes.c_int, [P, SZ, SZ, P, P]),
            "orc_lc_bind_top": (None, [P, SZ, F]),
            "orc_lc_round_msg_msb": (None, [P, P, SZ, P]),
            "orc_rs_fold_1b_rows": (None, [P, P, SZ, P]),
            "orc_rs_fold_b128_elems": (None, [P, P, SZ, P]),
            "orc_zc_lagrange_weights": (ctypes.c_int, [SZ, F, P]),
            "orc_zc_fold_at_z": (ctypes.c_int, [P, SZ, SZ, P, P]),
            "orc_zc_round_pair": (None, [P, P, SZ, P, P]),
            "orc_gf8_forward": (ctypes.c_int, [SZ, ctypes.c_uint8, P]),
            "orc_gf8_inverse": (ctypes.c_int, [SZ, ctypes.c_uint8, P]),
            "orc_zc_round1": (ctypes.c_int, [P, P, P, SZ, SZ, P, P, P]),
            "orc_zc_round1_mt": (ctypes.c_int, [P, P, P, SZ, SZ, P, P, P, SZ]),
            "orc_witness_layout": (None, [P]),
            "orc_blake3_block_witness": (None, [P, P, P, P]),
            "orc_blake3_witness": (ctypes.c_int, [P, SZ, SZ, P, P, P, P]),
            "orc_blake3_compress": (None, [P, P, U64, ctypes.c_uint32, ctypes.c_uint32, P]),
            "orc_pow_hash": (None, [P, U64, P]),
            "orc_blake3_block_witness_words": (None, [P, P, P, P]),
            "orc_blake3_witness_mt": (ctypes.c_int, [P, SZ, SZ, P, P, P, P, SZ]),
            "orc_blake3_xof": (None, [P, SZ, P, SZ]),
            "orc_merkle_multi_proof_indices": (SZ, [SZ, P, SZ, P]),
            "orc_merkle_multi_verify": (ctypes.c_int, [P, SZ, P, P, SZ, P, SZ]),
            "orc_gather_rows": (None, [P, SZ, P, SZ, P]),
            "orc_fs_new": (P, [P, SZ]),
            "orc_fs_free": (None, [P]),
            "orc_fs_label": (ctypes.c_int, [P, P, SZ]),
            "orc_fs_f128": (ctypes.c_int, [P, COrcF128]),
            "orc_fs_bytes": (ctypes.c_int, [P, P, SZ]),
            "orc_fs_f128_slice": (ctypes.c_int, [P, P, SZ]),
            "orc_fs_sample": (ctypes.c_int, [P, P]),
            "orc_fs_sample_vec": (ctypes.c_int, [P, SZ, P]),
            "orc_fs_digest": (None, [P, P]),
            "orc_rs_transpose": (None, [P, P]),
            "orc_sparse_fold": (None, [P, P, SZ, P, P, SZ]),
            "orc_rs_prove_batched": (ctypes.c_int, [P, P, SZ, P, SZ, P, P, P, P]),
            "orc_pcs_combined_basis_and_target": (ctypes.c_int, [P, P, SZ, P, SZ, P, P, P, P, P]),
            "orc_rs_claim_check": (COrcF128, [COrcF128, COrcF128, P]),
            "orc_ligerito_ood": (ctypes.c_int, [P, P, P, SZ, SZ, P, P, P]),
            "orc_leading_zero_bits_ok": (ctypes.c_int, [P, ctypes.c_uint32]),
            "orc_pow_ok": (ctypes.c_int, [P, U64, ctypes.c_uint32]),
            "orc_pow_grind": (U64, [P, ctypes.c_uint32]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        self.L = L

    # ---- scalars (ints) ----
    @staticmethod
    def _F(v):
        return COrcF128(v & (2**64 - 1), v >> 64)

    def mul(self, a, b):
        r = self.L.orc_f128_mul(self._F(a), self._F(b))
        return r.lo | (r.hi << 64)

    def inv(self, a):
        r = self.L.orc_f128_inv(self._F(a))
        return r.lo | (r.hi << 64)

    def mul_by_x(self, a):
        r = self.L.orc_f128_mul_by_x(self._F(a))
        return r.lo | (r.hi << 64)

    def clmul64(self, a, b, bitwise=False):
        lo, hi = ctypes.c_uint64(), ctypes.c_uint64()
        fn = self.L.orc_clmul64_bitwise if bitwise else self.L.orc_clmul64
        fn(a, b, ctypes.byref(lo), ctypes.byref(hi))
        return lo.value | (hi.value << 64)

    def subspace_poly_hat(self, i, z):
        r = self.L.orc_subspace_poly_hat(i, self._F(z))
        return r.lo | (r.hi << 64)

    def f8_mul(self, a, b):
        return self.L.orc_f8_mul(a, b)

    def phi8_table(self):
        t = np.zeros(256, dtype=F128_DT)
        self.L.orc_phi8_table(_p(t))
        return t

    # ---- arrays (numpy F128_DT) ----
    def fill(self, seed, n):
        a = np.zeros(n, dtype=F128_DT)
        self.L.orc_fill_f128(seed, _p(a), n)
        return a

    def mul_many(self, a, b):
        out = np.zeros(len(a), dtype=F128_DT)
        self.L.orc_f128_mul_many(_p(a), _p(b), _p(out), len(a))
        return out

    def twiddles(self, dim):
        t = np.zeros(max((1 << dim) - 1, 1), dtype=F128_DT)
        assert self.L.orc_ntt_twiddles(dim, _p(t)) == 0
        return t[: (1 << dim) - 1]

    def forward_interleaved(self, dim, data, num_ntts, start_layer=0):
        d = data.copy()
        assert self.L.orc_ntt_forward_interleaved(dim, _p(d), len(d), num_ntts, start_layer) == 0
        return d

    def inverse(self, dim, data):
        d = data.copy()
        assert self.L.orc_ntt_inverse(dim, _p(d), len(d)) == 0
        return d

    def rs_encode(self, dim, msg, n_cw, num_ntts):
        cw = np.zeros(n_cw, dtype=F128_DT)
        assert self.L.orc_ntt_rs_encode_interleaved(dim, _p(msg), len(msg), _p(cw), n_cw, num_ntts) == 0
        return cw

    # ---- BLAKE3 R1CS witness (oracle/witness.c) ----
    def witness_layout(self):
        out = np.zeros(6, dtype=np.uint64)
        self.L.orc_witness_layout(_p(out))
        return [int(x) for x in out]

    def block_witness(self, comp):
        """(z, a, b) of one compression record (COMPRESSION_DT scalar array of 1)."""
        c = np.ascontiguousarray(comp, dtype=COMPRESSION_DT).reshape(1)
        z, a, b = (np.zeros(256, np.uint64) for _ in range(3))
        self.L.orc_blake3_block_witness(_p(c), _p(z), _p(a), _p(b))
        return z, a, b

    def witness_mt(self, comps, n_blocks_log, threads, stripe=True):
        """Word-form restatement on `threads` POSIX threads (same outputs as witness)."""
        n_total = 1 << n_blocks_log
        c = np.ascontiguousarray(comps, dtype=COMPRESSION_DT)
        n = c.size
        if n == 0:
            c = np.zeros(1, COMPRESSION_DT)
        z, a, b = (np.zeros(n_total * 256, np.uint64) for _ in range(3))
        st = np.zeros(n_total * 2048, np.uint8) if stripe else None
        if self.L.orc_blake3_witness_mt(_p(c), n, n_blocks_log, _p(z), _p(a), _p(b), _p(st) if stripe else None,
                                        threads):
            raise ValueError("orc_blake3_witness_mt rejected the shape")
        return z, a, b, st

    def witness(self, comps, n_blocks_log, stripe=True):
        n_total = 1 << n_blocks_log
        c = np.ascontiguousarray(comps, dtype=COMPRESSION_DT)
        if c.size == 0:
            c = np.zeros(1, COMPRESSION_DT)
            n = 0
        else:
            n = c.size
        z, a, b = (np.zeros(n_total * 256, np.uint64) for _ in range(3))
        st = np.zeros(n_total * 2048, np.uint8) if stripe else None
        rc = self.L.orc_blake3_witness(_p(c), n, n_blocks_log, _p(z), _p(a), _p(b),
                                       _p(st) if stripe else None)
        if rc:
            raise ValueError("orc_blake3_witness rejected the shape")
        return z, a, b, st

    # ---- proof-of-work grinding (oracle/pow.c) ----
    def pow_hash(self, digest, nonce):
        d = np.frombuffer(bytes(digest), np.uint8).copy()
        out = np.zeros(32, np.uint8)
        self.L.orc_pow_hash(_p(d), nonce, _p(out))
        return out.tobytes()

    def lz_ok(self, h, bits):
        d = np.frombuffer(bytes(h), np.uint8).copy()
        return bool(self.L.orc_leading_zero_bits_ok(_p(d), bits))

    def pow_ok(self, digest, nonce, bits):
        d = np.frombuffer(bytes(digest), np.uint8).copy()
        return bool(self.L.orc_pow_ok(_p(d), nonce, bits))

    def pow_grind(self, digest, bits):
        d = np.frombuffer(bytes(digest), np.uint8).copy()
        return int(self.L.orc_pow_grind(_p(d), bits))

    def compress(self, cv, block_bytes, counter, block_len, flags):
        cvv = np.asarray(cv, np.uint32).copy()
        blk = np.frombuffer(bytes(block_bytes), np.uint8).copy()
        out = np.zeros(16, np.uint32)
        self.L.orc_blake3_compress(_p(cvv), _p(blk), counter, block_len, flags, _p(out))
        return out

    # ---- opening glue (oracle/open.c) ----
    def multi_proof_indices(self, num_leaves, positions):
        pos = np.asarray(positions, np.uint64).copy() if len(positions) else np.zeros(1, np.uint64)
        depth = max(num_leaves.bit_length() - 1, 1)
        out = np.zeros(max(len(positions) * depth, 1), np.uint64)
        k = self.L.orc_merkle_multi_proof_indices(num_leaves, _p(pos), len(positions), _p(out))
        return [int(x) for x in out[:k]]

    def multi_verify(self, root, num_leaves, sorted_positions, leaf_hashes, proof):
        pos = np.asarray(sorted_positions, np.uint64).copy() if len(sorted_positions) else np.zeros(1, np.uint64)
        lh = np.ascontiguousarray(leaf_hashes, np.uint8).reshape(-1) if len(sorted_positions) else np.zeros(32, np.uint8)
        pr = np.ascontiguousarray(proof, np.uint8).reshape(-1) if len(proof) else np.zeros(32, np.uint8)
        r = np.frombuffer(bytes(root), np.uint8).copy()
        return bool(self.L.orc_merkle_multi_verify(_p(r), num_leaves, _p(pos), _p(lh), len(sorted_positions),
                                                   _p(pr), len(proof)))

    def blake3_xof(self, data, n):
        b = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
        out = np.zeros(max(n, 1), np.uint8)
        self.L.orc_blake3_xof(_p(b), len(data), _p(out), n)
        return out[:n].tobytes()

    def blake3(self, data):
        b = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
        out = np.zeros(32, np.uint8)
        self.L.orc_blake3_hash(_p(b), len(data), _p(out))
        return out.tobytes()

    def leaf_cv(self, data):
        b = np.frombuffer(bytes(data), dtype=np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
        out = np.zeros(32, np.uint8)
        self.L.orc_blake3_leaf_cv(_p(b), len(data), _p(out))
        return out.tobytes()

    def parent_cv(self, l, r, root=False):
        a = np.frombuffer(l, np.uint8).copy()
        b = np.frombuffer(r, np.uint8).copy()
        out = np.zeros(32, np.uint8)
        (self.L.orc_blake3_parent_root if root else self.L.orc_blake3_parent_cv)(_p(a), _p(b), _p(out))
        return out.tobytes()

    def merkle_tree(self, data_bytes, num_leaves):
        d = np.frombuffer(data_bytes, np.uint8).copy() if isinstance(data_bytes, (bytes, bytearray)) else data_bytes.view(np.uint8).reshape(-1).copy()
        tree = np.zeros((2 * num_leaves - 1, 32), np.uint8)
        assert self.L.orc_merkle_tree(_p(d), d.size, num_leaves, _p(tree)) == 0
        return tree

    def merkle_proof(self, tree, num_leaves, index):
        path = np.zeros((max(num_leaves.bit_length() - 1, 1), 32), np.uint8)
        k = self.L.orc_merkle_proof(_p(tree), num_leaves, index, _p(path))
        return path[:k]

    def merkle_verify(self, root, leaf_hash, index, path):
        r = np.frombuffer(bytes(root), np.uint8).copy()
        lh = np.frombuffer(bytes(leaf_hash), np.uint8).copy()
        p = np.ascontiguousarray(path) if len(path) else np.zeros((1, 32), np.uint8)
        return bool(self.L.orc_merkle_verify(_p(r), _p(lh), index, _p(p), len(path)))

    def commit(self, msg, num_ntts, log_inv_rate):
        n_cw = len(msg) << log_inv_rate
        n_pos = n_cw // num_ntts
        cw = np.zeros(n_cw, dtype=F128_DT)
        tree = np.zeros((2 * n_pos - 1, 32), np.uint8)
        root = np.zeros(32, np.uint8)
        assert self.L.orc_commit(_p(msg), len(msg), num_ntts, log_inv_rate, _p(cw), _p(tree), _p(root)) == 0
        return cw, tree, root.tobytes()

    def commit_mt(self, msg, num_ntts, log_inv_rate, n_threads):
        n_cw = len(msg) << log_inv_rate
        n_pos = n_cw // num_ntts
        cw = np.zeros(n_cw, dtype=F128_DT)
        tree = np.zeros((2 * n_pos - 1, 32), np.uint8)
        root = np.zeros(32, np.uint8)
        assert self.L.orc_commit_mt(_p(msg), len(msg), num_ntts, log_inv_rate, _p(cw), _p(tree), _p(root),
                                    n_threads) == 0
        return cw, tree, root.tobytes()

    def build_eq(self, point):
        p = ints_to_f128(point) if not isinstance(point, np.ndarray) else point
        out = np.zeros(1 << len(p), dtype=F128_DT)
        self.L.orc_build_eq(_p(p) if len(p) else None, len(p), _p(out))
        return out

    def partial_eval_lsb(self, evals, rs):
        r = ints_to_f128(rs)
        out = np.zeros(len(evals) >> len(rs), dtype=F128_DT)
        self.L.orc_partial_eval_lsb(_p(evals), len(evals), _p(r) if len(rs) else None, len(rs), _p(out))
        return out

    def round_msg(self, f, b):
        out = np.zeros(2, dtype=F128_DT)
        self.L.orc_sc_round_msg(_p(f), _p(b), len(f), _p(out))
        return out

    def round_msg_and_eval(self, f, b):
        out = np.zeros(3, dtype=F128_DT)
        self.L.orc_sc_round_msg_and_eval(_p(f), _p(b), len(f), _p(out))
        return out

    def fold_and_msg(self, f, b, r):
        h = len(f) // 2
        fo = np.zeros(h, dtype=F128_DT)
        bo = np.zeros(h, dtype=F128_DT)
        out = np.zeros(2, dtype=F128_DT)
        self.L.orc_sc_fold_and_msg(_p(f), _p(b), len(f), self._F(r), _p(fo), _p(bo), _p(out))
        return fo, bo, out

    def glue(self, b, b_new, alpha):
        bb = b.copy()
        self.L.orc_sc_glue(_p(bb), _p(b_new), len(b), self._F(alpha))
        return bb

    def eval_sk_at_vks(self, log_n):
        out = np.zeros(log_n + 1, dtype=F128_DT)
        self.L.orc_eval_sk_at_vks(log_n, _p(out))
        return out

    def induce(self, log_msg_cols, rows, v_challenges, queries, alpha):
        nq, row_len = rows.shape
        v = ints_to_f128(v_challenges)
        a = ints_to_f128(alpha)
        q = np.asarray(queries, dtype=np.uint64)
        basis = np.zeros(1 << log_msg_cols, dtype=F128_DT)
        s = np.zeros(1, dtype=F128_DT)
        rr = np.ascontiguousarray(rows)
        assert self.L.orc_induce_sumcheck_poly(log_msg_cols, _p(rr), nq, row_len, _p(v) if len(v) else None, len(v),
                                               _p(q), _p(a) if len(a) else None, len(a), _p(basis), _p(s)) == 0
        return basis, f128_to_ints(s)[0]

    # ---- lincheck / ring switch / zerocheck ----
    def lc_partial_fold(self, z_bytes, m, k_log, eq_outer):
        z = np.ascontiguousarray(z_bytes, dtype=np.uint8)
        out = np.zeros(1 << k_log, dtype=F128_DT)
        assert self.L.orc_lc_partial_fold(_p(z), m, k_log, _p(eq_outer), _p(out)) == 0
        return out

    def lc_bind_top(self, v, r):
        vv = v.copy()
        self.L.orc_lc_bind_top(_p(vv), len(vv), self._F(r))
        return vv[: len(vv) // 2]

    def lc_round_msg_msb(self, comb, z):
        out = np.zeros(2, dtype=F128_DT)
        self.L.orc_lc_round_msg_msb(_p(comb), _p(z), len(comb), _p(out))
        return out

    def rs_fold_1b_rows(self, z, eq):
        out = np.zeros(128, dtype=F128_DT)
        self.L.orc_rs_fold_1b_rows(_p(z), _p(eq), len(z), _p(out))
        return out

    def rs_fold_b128_elems(self, suffix, eq_r):
        out = np.zeros(len(suffix), dtype=F128_DT)
        self.L.orc_rs_fold_b128_elems(_p(suffix), _p(eq_r), len(suffix), _p(out))
        return out

    # ---- ring switch proving and claim combination (ring_switch.c) ----
    def fs(self, domain):
        return OracleFs(self, domain)

    def sparse_fold(self, ptr, cols, w, n_cols):
        """out[c] = XOR of w[r] over the nonzeros (r, c) of the CSR matrix (ptr, cols)."""
        p = np.ascontiguousarray(ptr, dtype=np.uint32)
        c = np.ascontiguousarray(cols, dtype=np.uint32)
        out = np.zeros(n_cols, dtype=F128_DT)
        self.L.orc_sparse_fold(_p(p), _p(c), len(p) - 1, _p(w), _p(out), n_cols)
        return out

    def rs_transpose(self, s_hat_v):
        a = ints_to_f128(s_hat_v) if not isinstance(s_hat_v, np.ndarray) else s_hat_v
        out = np.zeros(128, dtype=F128_DT)
        self.L.orc_rs_transpose(_p(a), _p(out))
        return f128_to_ints(out)

    def _rs_args(self, z, x_outers):
        log_len = int(len(z)).bit_length() - 1
        assert len(z) == 1 << log_len and all(len(x) == log_len + 1 for x in x_outers)
        xs = ints_to_f128([v for x in x_outers for v in x])
        k = len(x_outers)
        return log_len, xs, k, np.zeros(128 * k, F128_DT), np.zeros(k, F128_DT), np.zeros(k, F128_DT)

    def rs_prove_batched(self, fs, z, x_outers, basis=True):
        """(s_hat_vs, claims, gammas, b_combined or None) of ring_switch::prove_batched."""
        log_len, xs, k, shv, cl, ga = self._rs_args(z, x_outers)
        b = np.zeros(1 << log_len, F128_DT) if basis else None
        assert self.L.orc_rs_prove_batched(fs.h, _p(z), log_len, _p(xs), k, _p(shv), _p(cl), _p(ga),
                                           _p(b) if basis else None) == 0
        v = f128_to_ints(shv)
        return [v[128 * j:128 * (j + 1)] for j in range(k)], f128_to_ints(cl), f128_to_ints(ga), b

    def pcs_combined_basis_and_target(self, fs, z, x_outers):
        """(s_hat_vs, claims, gammas, b_combined, target) of compute_combined_basis_and_target."""
        log_len, xs, k, shv, cl, ga = self._rs_args(z, x_outers)
        b = np.zeros(1 << log_len, F128_DT)
        t = np.zeros(1, F128_DT)
        assert self.L.orc_pcs_combined_basis_and_target(fs.h, _p(z), log_len, _p(xs), k, _p(shv), _p(cl), _p(ga),
                                                        _p(b), _p(t)) == 0
        v = f128_to_ints(shv)
        return [v[128 * j:128 * (j + 1)] for j in range(k)], f128_to_ints(cl), f128_to_ints(ga), b, \
            f128_to_ints(t)[0]

    def ligerito_ood(self, fs, f, b, n_samples, t_r):
        """The OOD loop: returns (b', T', ood_values, msgs as (u_0, u_2) pairs); b is not modified."""
        log_n = int(len(f)).bit_length() - 1
        b2 = b.copy()
        t = ints_to_f128([t_r])
        ys = np.zeros(max(n_samples, 1), F128_DT)
        ms = np.zeros(2 * max(n_samples, 1), F128_DT)
        assert self.L.orc_ligerito_ood(fs.h, _p(f), _p(b2), log_n, n_samples, _p(t), _p(ys), _p(ms)) == 0
        m = f128_to_ints(ms)
        return b2, f128_to_ints(t)[0], f128_to_ints(ys)[:n_samples], [(m[2 * i], m[2 * i + 1])
                                                                      for i in range(n_samples)]

    def rs_claim_check(self, z_skip, x_outer_0, s_hat_v):
        a = ints_to_f128(s_hat_v)
        r = self.L.orc_rs_claim_check(self._F(z_skip), self._F(x_outer_0), _p(a))
        return r.lo | (r.hi << 64)

    def zc_lagrange_weights(self, k_skip, z):
        out = np.zeros(1 << k_skip, dtype=F128_DT)
        assert self.L.orc_zc_lagrange_weights(k_skip, self._F(z), _p(out)) == 0
        return out

    def zc_fold_at_z(self, bits_bytes, m, k_skip, weights):
        b = np.ascontiguousarray(bits_bytes, dtype=np.uint8)
        out = np.zeros(1 << (m - k_skip), dtype=F128_DT)
        assert self.L.orc_zc_fold_at_z(_p(b), m, k_skip, _p(weights), _p(out)) == 0
        return out

    def zc_round_pair(self, a, b, eq_rest):
        out = np.zeros(2, dtype=F128_DT)
        self.L.orc_zc_round_pair(_p(a), _p(b), len(a), _p(eq_rest), _p(out))
        return out

    def gf8_forward(self, k, beta, v):
        a = np.array(v, dtype=np.uint8)
        assert self.L.orc_gf8_forward(k, beta, _p(a)) == 0
        return [int(x) for x in a]

    def gf8_inverse(self, k, beta, v):
        a = np.array(v, dtype=np.uint8)
        assert self.L.orc_gf8_inverse(k, beta, _p(a)) == 0
        return [int(x) for x in a]

    def zc_round1(self, a, b, c, m, k_skip, r):
        rr = ints_to_f128(r)
        pab = np.zeros(1 << k_skip, dtype=F128_DT)
        pc = np.zeros(1 << k_skip, dtype=F128_DT)
        args = [np.ascontiguousarray(x, dtype=np.uint8) for x in (a, b, c)]
        assert self.L.orc_zc_round1(_p(args[0]), _p(args[1]), _p(args[2]), m, k_skip, _p(rr), _p(pab), _p(pc)) == 0
        return pab, pc

    def zc_round1_mt(self, a, b, c, m, k_skip, r, threads):
        """orc_zc_round1 on `threads` threads (k_skip = 6; same outputs)."""
        rr = ints_to_f128(r)
        pab = np.zeros(1 << k_skip, dtype=F128_DT)
        pc = np.zeros(1 << k_skip, dtype=F128_DT)
        args = [np.ascontiguousarray(x, dtype=np.uint8) for x in (a, b, c)]
        assert self.L.orc_zc_round1_mt(_p(args[0]), _p(args[1]), _p(args[2]), m, k_skip, _p(rr), _p(pab), _p(pc),
                                       threads) == 0
        return pab, pc


class OracleFs:
    """The oracle's Fiat-Shamir transcript (a BLAKE3 byte log, ring_switch.c)."""

    def __init__(self, orc, domain):
        d = bytes(domain)
        self.orc = orc
        self.h = orc.L.orc_fs_new(d or None, len(d))
        assert self.h

    def __del__(self):
        if getattr(self, "h", None):
            self.orc.L.orc_fs_free(self.h)
            self.h = None

    def observe_label(self, label):
        b = bytes(label)
        assert self.orc.L.orc_fs_label(self.h, b or None, len(b)) == 0

    def observe_f128(self, v):
        assert self.orc.L.orc_fs_f128(self.h, Oracle._F(v)) == 0

    def observe_f128_slice(self, vals):
        a = ints_to_f128(vals)
        assert self.orc.L.orc_fs_f128_slice(self.h, _p(a) if len(vals) else None, len(vals)) == 0

    def sample_f128(self):
        out = np.zeros(1, F128_DT)
        assert self.orc.L.orc_fs_sample(self.h, _p(out)) == 0
        return f128_to_ints(out)[0]

    def sample_f128_vec(self, n):
        out = np.zeros(max(n, 1), F128_DT)
        assert self.orc.L.orc_fs_sample_vec(self.h, n, _p(out)) == 0
        return f128_to_ints(out)[:n]

    def state_digest(self):
        out = np.zeros(32, np.uint8)
        self.orc.L.orc_fs_digest(self.h, _p(out))
        return out.tobytes()

    def observe_bytes(self, data):
        b = bytes(data)
        assert self.orc.L.orc_fs_bytes(self.h, b or None, len(b)) == 0

    def verify_pow(self, nonce, bits):
        """verify_pow (CF/challenger.rs:549-601): check against the digest, then absorb u64le(nonce)."""
        ok = self.orc.pow_ok(self.state_digest(), nonce, bits)
        self.observe_bytes(int(nonce).to_bytes(8, "little"))
        return ok

    def sample_distinct_queries(self, block_len, count):
        """sample_distinct_queries (CF/pcs/ligerito.rs:9665-9720): lo mod block_len, first
        occurrence kept, sorted."""
        seen, out = set(), []
        while len(out) < count:
            q = (self.sample_f128() & (2**64 - 1)) % block_len
            if q not in seen:
                seen.add(q)
                out.append(q)
        return sorted(out)
