"""Python mirror of the reference's hot-path interface over the device library.

Names, argument meaning and error behaviour follow the reference entry points
(`CF/` = crates/flock-core/src/):

  AdditiveNttF128.standard / twiddle / rs_encode_interleaved /
      forward_transform_interleaved        CF/ntt/additive_ntt_f128.rs:2661-2830
  merkle_tree / hash_leaves / hash_pairs_level  CF/merkle.rs:498,683,753
  PcsParams / commit                       CF/pcs/commit.rs:34-140,187
  ligero_commit                            CF/pcs/ligerito.rs:4750
  build_eq_table                           CF/lincheck.rs:2080
  partial_eval_lsb                         CF/pcs/ligerito.rs:1162
  SumcheckProver (new/fold/introduce_new/glue)  CF/pcs/ligerito.rs:8893-9260
  generate_witness_with_ab_packed[_and_lincheck]
                                           CP/r1cs_hashes/blake3.rs:2184,4617
                                           (CP = crates/flock-prover/src/)
  pow_scan / grind_pow_nonce               CF/challenger.rs:384-474,1648
  FsChallenger (BLAKE3 transcript)         CF/challenger.rs:226-523
  gather_rows / merkle_multi_proof         CF/merkle.rs:1020-1107 (query openings)

Buffers are torch tensors on the GPU: F128 arrays are int64 tensors of shape
(n, 2) holding (lo, hi); hashes are uint8 tensors of shape (n, 32).  Work is
enqueued on torch's current stream.  Shape errors raise ValueError before any
launch; device failures raise FlockGpuError (and latch the library off, as the
reference's GPU facade does, CF/gpu.rs:455-466).
"""
from dataclasses import dataclass

import torch

from . import _native as nat
from ._native import F128, FlockGpuError  # noqa: F401  (re-exported)


import ctypes as _ctypes  # noqa: E402

ctypes_u64 = _ctypes.c_uint64


def _stream():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(v):
    import ctypes
    return ctypes.c_void_p(v)


def _ptr(t):
    import ctypes
    return ctypes.c_void_p(t.data_ptr())


def _check_f128(t, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.int64 and t.dim() == 2
            and t.shape[1] == 2 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous CUDA int64 tensor of shape (n, 2)")
    return t.shape[0]


def f128_empty(n, device="cuda"):
    return torch.empty((n, 2), dtype=torch.int64, device=device)


def hashes_empty(n, device="cuda"):
    return torch.empty((n, 32), dtype=torch.uint8, device=device)


def _pow2(n):
    return n > 0 and (n & (n - 1)) == 0


def _log2(n):
    return n.bit_length() - 1


class AdditiveNttF128:
    """AdditiveNttF128::standard(dim): basis {1, x, ..., x^(dim-1)}."""

    def __init__(self, dim):
        if not 0 <= dim <= 24:
            raise ValueError("standard NTT: dim must be in 0..24 on the device")
        self.dim = dim
        self._tw = None

    @classmethod
    def standard(cls, dim):
        return cls(dim)

    def log_domain_size(self):
        return self.dim

    def twiddles(self):
        """Breadth-first twiddle table (2^dim - 1 values, layer l at 2^l - 1) as ints."""
        if self._tw is None:
            n = (1 << self.dim) - 1
            buf = (F128 * max(n, 1))()
            nat.call("flk_ntt_twiddles", self.dim, buf)
            self._tw = [buf[i].to_int() for i in range(n)]
        return self._tw

    def twiddle(self, layer, block):
        if not (layer < self.dim and block < (1 << layer)):
            raise ValueError("twiddle: (layer, block) out of range")
        return self.twiddles()[(1 << layer) - 1 + block]

    def forward_transform_interleaved(self, data, num_ntts, start_layer=0):
        n = _check_f128(data, "data")
        if not _pow2(num_ntts) or n % num_ntts or not _pow2(n // num_ntts):
            raise ValueError("forward_transform_interleaved: bad shape")
        if _log2(n // num_ntts) > self.dim or start_layer > _log2(n // num_ntts):
            raise ValueError("forward_transform_interleaved: domain exceeds the NTT")
        nat.call("flk_ntt_forward_interleaved", self.dim, _ptr(data), n, num_ntts, start_layer, _stream())

    def forward_transform(self, data):
        """AdditiveNttF128::forward_transform (:2701): one transform, in place."""
        self.forward_transform_interleaved(data, 1)

    def inverse_transform_interleaved(self, data, num_ntts):
        """Inverse of forward_transform_interleaved, in place (inverse_transform, :6322)."""
        n = _check_f128(data, "data")
        if not _pow2(num_ntts) or n % num_ntts or not _pow2(n // num_ntts):
            raise ValueError("inverse_transform_interleaved: bad shape")
        if _log2(n // num_ntts) > self.dim:
            raise ValueError("inverse_transform_interleaved: domain exceeds the NTT")
        nat.call("flk_ntt_inverse_interleaved", self.dim, _ptr(data), n, num_ntts, _stream())

    def inverse_transform(self, data):
        """AdditiveNttF128::inverse_transform (:6322-6342): exact inverse of forward_transform."""
        self.inverse_transform_interleaved(data, 1)

    def rs_encode_interleaved(self, msg, codeword, num_ntts):
        n_msg = _check_f128(msg, "msg")
        n_cw = _check_f128(codeword, "codeword")
        if not _pow2(num_ntts) or n_msg == 0 or n_msg % num_ntts or n_cw % n_msg:
            raise ValueError("rs_encode_interleaved: bad shape")
        inv_rate = n_cw // n_msg
        if not _pow2(inv_rate) or inv_rate < 2:
            raise ValueError("rs_encode_interleaved: inverse rate must be a power of two >= 2")
        if _log2(n_cw // num_ntts) > self.dim:
            raise ValueError("rs_encode_interleaved: domain exceeds the NTT")
        nat.call("flk_ntt_rs_encode_interleaved", self.dim, _ptr(msg), n_msg, _ptr(codeword), n_cw, num_ntts,
                 _stream())


def hash_leaves(data, leaf_size, out):
    """merkle::hash_leaves (BLAKE3): out[i] = non-root CV of leaf i."""
    n = out.shape[0]
    if data.numel() * data.element_size() != n * leaf_size:
        raise ValueError("hash_leaves: data length != n_leaves * leaf_size")
    nat.call("flk_merkle_hash_leaves", _ptr(data), leaf_size, n, _ptr(out), _stream())


def hash_pairs_level(children, out):
    n_children = children.shape[0]
    if out.shape[0] * 2 != n_children:
        raise ValueError("hash_pairs_level: output must have half as many nodes")
    nat.call("flk_merkle_hash_pairs_level", _ptr(children), n_children, _ptr(out), _stream())


def merkle_tree(data, num_leaves, kind=nat.FLK_HASH_BLAKE3, tree=None):
    """merkle::merkle_tree: flat 2n-1 tree (leaves first, root last)."""
    nbytes = data.numel() * data.element_size()
    if not _pow2(num_leaves) or nbytes % num_leaves:
        raise ValueError("merkle_tree: num_leaves must be a power of two dividing the data length")
    if tree is None:
        tree = hashes_empty(2 * num_leaves - 1, data.device)
    nat.call("flk_merkle_tree", _ptr(data), nbytes, num_leaves, kind, _ptr(tree), _stream())
    return tree


SIZE_MAX = (1 << 64) - 1


def _host_or_device_ptr(buf):
    """Address of a torch tensor (CUDA or CPU) or a numpy array."""
    import ctypes
    if isinstance(buf, torch.Tensor):
        if not buf.is_contiguous():
            raise ValueError("buffers must be contiguous")
        return ctypes.c_void_p(buf.data_ptr()), buf.numel() * buf.element_size()
    import numpy as np
    if isinstance(buf, np.ndarray):
        if not buf.flags["C_CONTIGUOUS"]:
            raise ValueError("buffers must be contiguous")
        return ctypes.c_void_p(buf.ctypes.data), buf.nbytes
    raise TypeError("expected a torch tensor or a numpy array")


def prewire(data):
    """crate::gpu::prewire (CF/gpu.rs:497-515): accepted; no host registration."""
    p, n = _host_or_device_ptr(data)
    nat.call("flk_prewire", p, n)


class MerkleSession:
    """crate::gpu::merkle::Session (CF/gpu.rs:1228-1490): leaves hashed in
    committed ranges, then the parent levels with >= stop_nodes nodes; data and
    tree may live on the host (numpy / CPU tensor) or the device."""

    def __init__(self, handle, data, tree):
        self._h = handle
        self._keep = (data, tree)  # the buffers must outlive the session

    @staticmethod
    def begin(data, leaf_size, tree, stop_nodes=SIZE_MAX):
        import ctypes
        dp, dn = _host_or_device_ptr(data)
        tp, tn = _host_or_device_ptr(tree)
        h = ctypes.c_void_p()
        nat.call("flk_merkle_begin", dp, dn, leaf_size, tp, tn // 32, stop_nodes, _stream(), ctypes.byref(h))
        return MerkleSession(h, data, tree)

    def n_leaves(self):
        return nat.load().flk_merkle_session_leaves(self._h)

    def commit_leaves(self, lo, hi):
        nat.call("flk_merkle_commit_leaves", self._h, lo, hi)

    def commit_parent_levels(self):
        nat.call("flk_merkle_commit_parent_levels", self._h)

    def finish(self):
        """Waits; returns the session's device time in seconds."""
        import ctypes
        sec = ctypes.c_double(0.0)
        h, self._h = self._h, None
        nat.call("flk_merkle_finish", h, ctypes.byref(sec))
        return sec.value

    def __del__(self):
        if getattr(self, "_h", None):
            nat.load().flk_merkle_discard(self._h)
            self._h = None


@dataclass
class PcsParams:
    """CF/pcs/commit.rs:34-95 (LOG_PACKING = 7)."""

    m: int
    log_inv_rate: int = 1
    log_batch_size: int = 6
    merkle_hash: int = nat.FLK_HASH_BLAKE3

    def log_msg_len(self):
        return self.m - 7

    def log_dim(self):
        return self.log_msg_len() - self.log_batch_size

    def k_code(self):
        return self.log_dim() + self.log_inv_rate

    def n_positions(self):
        return 1 << self.k_code()

    def num_ntts(self):
        return 1 << self.log_batch_size

    def codeword_len_f128(self):
        return self.n_positions() * self.num_ntts()

    def n_leaves(self):
        return self.n_positions()

    def leaf_size_bytes(self):
        return 16 << self.log_batch_size


@dataclass
class ProverData:
    codeword: torch.Tensor
    merkle_tree: torch.Tensor

    def root(self):
        return self.merkle_tree[-1]


def ligero_commit(poly, log_msg_cols, log_num_interleaved, log_inv_rate, kind=nat.FLK_HASH_BLAKE3,
                  codeword=None, tree=None):
    """ligerito::ligero_commit: RS-encode the interleaved matrix, Merkle-hash its rows."""
    n = _check_f128(poly, "poly")
    if n != 1 << (log_msg_cols + log_num_interleaved):
        raise ValueError("ligero_commit: poly length != 2^(log_msg_cols + log_num_interleaved)")
    n_pos = 1 << (log_msg_cols + log_inv_rate)
    if codeword is None:
        codeword = f128_empty(n_pos << log_num_interleaved, poly.device)
    if tree is None:
        tree = hashes_empty(2 * n_pos - 1, poly.device)
    nat.call("flk_commit", _ptr(poly), log_msg_cols, log_num_interleaved, log_inv_rate, kind, _ptr(codeword),
             _ptr(tree), None, _stream())
    return ProverData(codeword, tree)


def commit(z_packed, params):
    """pcs::commit: returns (root tensor, ProverData)."""
    n = _check_f128(z_packed, "z_packed")
    if n != 1 << params.log_msg_len():
        raise ValueError("commit: z_packed length must be 2^(m-7)")
    if params.log_inv_rate < 1:
        raise ValueError("log_inv_rate must be >= 1 for a non-trivial RS code")
    pd = ligero_commit(z_packed, params.log_dim(), params.log_batch_size, params.log_inv_rate, params.merkle_hash)
    return pd.root(), pd


def build_eq_table(point, out=None, device="cuda"):
    """lincheck::build_eq_table: LSB-first eq table of `point` (list of ints)."""
    d = len(point)
    if out is None:
        out = f128_empty(1 << d, device)
    arr = (F128 * max(d, 1))(*[F128.from_int(v) for v in point])
    nat.call("flk_build_eq", arr, d, _ptr(out), _stream())
    return out


def partial_eval_lsb(evals, rs, out=None):
    n = _check_f128(evals, "evals")
    k = len(rs)
    if not _pow2(n) or (1 << k) > n:
        raise ValueError("partial_eval_lsb: need 2^k <= len(evals), power of two")
    if out is None:
        out = f128_empty(n >> k, evals.device)
    arr = (F128 * max(k, 1))(*[F128.from_int(v) for v in rs])
    nat.call("flk_partial_eval_lsb", _ptr(evals), n, arr, k, _ptr(out), _stream())
    return out


def round_msg(f, b, out=None):
    n = _check_f128(f, "f")
    if _check_f128(b, "b") != n or not _pow2(n) or n < 2:
        raise ValueError("round_msg: f and b must have equal power-of-two length >= 2")
    if out is None:
        out = f128_empty(2, f.device)
    nat.call("flk_sc_round_msg", _ptr(f), _ptr(b), n, _ptr(out), _stream())
    return out


def round_msg_and_eval(f, b, out=None):
    n = _check_f128(f, "f")
    if _check_f128(b, "b") != n or not _pow2(n) or n < 2:
        raise ValueError("round_msg_and_eval: f and b must have equal power-of-two length >= 2")
    if out is None:
        out = f128_empty(3, f.device)
    nat.call("flk_sc_round_msg_and_eval", _ptr(f), _ptr(b), n, _ptr(out), _stream())
    return out


def fold_and_msg(f, b, r, f_out=None, b_out=None, out=None):
    n = _check_f128(f, "f")
    if _check_f128(b, "b") != n or not _pow2(n) or n < 4:
        raise ValueError("fold_and_msg: f and b must have equal power-of-two length >= 4")
    f_out = f128_empty(n // 2, f.device) if f_out is None else f_out
    b_out = f128_empty(n // 2, f.device) if b_out is None else b_out
    out = f128_empty(2, f.device) if out is None else out
    nat.call("flk_sc_fold_and_msg", _ptr(f), _ptr(b), n, F128.from_int(r), _ptr(f_out), _ptr(b_out), _ptr(out),
             _stream())
    return f_out, b_out, out


def glue(b, b_new, alpha):
    n = _check_f128(b, "b")
    if _check_f128(b_new, "b_new") != n:
        raise ValueError("glue: length mismatch")
    nat.call("flk_sc_glue", _ptr(b), _ptr(b_new), n, F128.from_int(alpha), _stream())


def induce_sumcheck_poly(log_msg_cols, opened_rows, v_challenges, queries, alpha, basis=None, enforced_sum=None):
    """ligerito::induce_sumcheck_poly: (basis_poly, enforced_sum) on the device.

    opened_rows: (n_queries * 2^len(v), 2) int64 tensor, row-major; queries: CUDA
    int64 tensor of query indices; v_challenges, alpha: lists of ints."""
    import ctypes
    nq = int(queries.numel())
    n_v = len(v_challenges)
    if _check_f128(opened_rows, "opened_rows") != nq << n_v:
        raise ValueError("induce_sumcheck_poly: rows must be n_queries x 2^len(v_challenges)")
    if nq > (1 << len(alpha)):
        raise ValueError("induce_sumcheck_poly: alpha too short for the query count")
    basis = f128_empty(1 << log_msg_cols, opened_rows.device) if basis is None else basis
    enforced_sum = f128_empty(1, opened_rows.device) if enforced_sum is None else enforced_sum
    v = (F128 * max(n_v, 1))(*[F128.from_int(x) for x in v_challenges])
    a = (F128 * max(len(alpha), 1))(*[F128.from_int(x) for x in alpha])
    nat.call("flk_induce_sumcheck_poly", log_msg_cols, _ptr(opened_rows), nq, v, n_v, _ptr(queries), a, len(alpha),
             _ptr(basis), _ptr(enforced_sum), _stream())
    return basis, enforced_sum


# ---- lincheck (CF/lincheck.rs) -------------------------------------------------
def partial_fold_packed_z(z_packed, m, k_log, eq_outer, out=None):
    """lincheck::partial_fold_packed_z: z_packed is a CUDA uint8 tensor in the stripe layout."""
    if z_packed.numel() != (1 << m) // 8:
        raise ValueError("partial_fold_packed_z: z_packed must hold 2^m bits")
    if _check_f128(eq_outer, "eq_outer") != 1 << (m - k_log):
        raise ValueError("partial_fold_packed_z: eq_outer must have 2^(m-k_log) entries")
    out = f128_empty(1 << k_log, z_packed.device) if out is None else out
    nat.call("flk_lc_partial_fold", _ptr(z_packed), m, k_log, _ptr(eq_outer), _ptr(out), _stream())
    return out


def lincheck_round_msg(comb, z, out=None):
    """(q(1), q(inf)) of the product sumcheck over the top variable."""
    n = _check_f128(comb, "comb")
    if _check_f128(z, "z") != n or not _pow2(n) or n < 2:
        raise ValueError("lincheck_round_msg: equal power-of-two lengths >= 2 required")
    out = f128_empty(2, comb.device) if out is None else out
    nat.call("flk_lc_round_msg_msb", _ptr(comb), _ptr(z), n, _ptr(out), _stream())
    return out


def sumcheck_bind_top_in_place(v, r):
    """Binds the top variable in place; returns the bound half as a view."""
    n = _check_f128(v, "v")
    if not _pow2(n) or n < 2:
        raise ValueError("sumcheck_bind_top_in_place: power-of-two length >= 2 required")
    nat.call("flk_lc_bind_top", _ptr(v), n, F128.from_int(r), _stream())
    return v[: n // 2]


# ---- ring switching (CF/pcs/ring_switch.rs) -------------------------------------
def fold_1b_rows(packed_witness, suffix_tensor, out=None):
    n = _check_f128(packed_witness, "packed_witness")
    if _check_f128(suffix_tensor, "suffix_tensor") != n:
        raise ValueError("fold_1b_rows: length mismatch")
    out = f128_empty(128, packed_witness.device) if out is None else out
    nat.call("flk_rs_fold_1b_rows", _ptr(packed_witness), _ptr(suffix_tensor), n, _ptr(out), _stream())
    return out


def fold_b128_elems(suffix_tensor, eq_r_dprime, out=None):
    n = _check_f128(suffix_tensor, "suffix_tensor")
    if _check_f128(eq_r_dprime, "eq_r_dprime") != 128:
        raise ValueError("fold_b128_elems: eq_r_dprime must have 128 entries")
    out = f128_empty(n, suffix_tensor.device) if out is None else out
    nat.call("flk_rs_fold_b128_elems", _ptr(suffix_tensor), _ptr(eq_r_dprime), n, _ptr(out), _stream())
    return out


def _rs_claims(packed_witness, x_outers):
    n = _check_f128(packed_witness, "packed_witness")
    if not _pow2(n):
        raise ValueError("ring switch: packed witness length must be a power of two")
    log_len = _log2(n)
    x_outers = [list(x) for x in x_outers]
    if not x_outers:
        raise ValueError("ring switch: at least one claim is required")
    for x in x_outers:
        if len(x) != log_len + 1:
            raise ValueError(f"ring switch: x_outer must have {log_len + 1} coordinates (2^L = len(packed_witness))")
    flat = [v for x in x_outers for v in x]
    arr = (F128 * len(flat))(*[F128.from_int(v) for v in flat])
    return log_len, arr, len(x_outers)


def _rs_outputs(k):
    return (F128 * (128 * k))(), (F128 * k)(), (F128 * k)()


def prove_batched(packed_witness, x_outers, challenger, b_combined=None, want_basis=True):
    """ring_switch::prove_batched (CF/pcs/ring_switch.rs:3727-3735, 4310-4383):
    returns (s_hat_vs, sumcheck_claims, gammas, b_combined), where b_combined =
    sum_k of the gamma-baked rs_eq_ind (a device tensor, or None when
    want_basis is False)."""
    log_len, arr, k = _rs_claims(packed_witness, x_outers)
    if want_basis and b_combined is None:
        b_combined = f128_empty(1 << log_len, packed_witness.device)
    shv, cl, ga = _rs_outputs(k)
    nat.call("flk_rs_prove_batched", challenger._h, _ptr(packed_witness), log_len, arr, k, shv, cl, ga,
             _ptr(b_combined) if want_basis else None, _stream())
    s_hat_vs = [[shv[128 * j + r].to_int() for r in range(128)] for j in range(k)]
    return s_hat_vs, [c.to_int() for c in cl], [g.to_int() for g in ga], (b_combined if want_basis else None)


def compute_combined_basis_and_target(packed_witness, x_outers, challenger, b_combined=None):
    """compute_combined_basis_and_target (CF/pcs.rs:778-870) for ring-switched
    claims: returns (s_hat_vs, sumcheck_claims, gammas, b_combined, target)."""
    log_len, arr, k = _rs_claims(packed_witness, x_outers)
    if b_combined is None:
        b_combined = f128_empty(1 << log_len, packed_witness.device)
    shv, cl, ga = _rs_outputs(k)
    tgt = F128()
    import ctypes
    nat.call("flk_pcs_combined_basis_and_target", challenger._h, _ptr(packed_witness), log_len, arr, k, shv, cl, ga,
             _ptr(b_combined), ctypes.byref(tgt), _stream())
    s_hat_vs = [[shv[128 * j + r].to_int() for r in range(128)] for j in range(k)]
    return s_hat_vs, [c.to_int() for c in cl], [g.to_int() for g in ga], b_combined, tgt.to_int()


# ---- zerocheck (CF/zerocheck) ----------------------------------------------------
def lagrange_weights(k_skip, z):
    """lagrange_weights_naive on the phi8 nodes (host), as a list of ints."""
    buf = (F128 * (1 << k_skip))()
    nat.call("flk_zc_lagrange_weights", k_skip, F128.from_int(z), buf)
    return [buf[i].to_int() for i in range(1 << k_skip)]


def fold_at_z(bits, m, k_skip, weights, out=None):
    """fold_at_z_naive: bits is a CUDA uint8 tensor of 2^m witness bits; weights a (2^k, 2) tensor."""
    if bits.numel() != (1 << m) // 8:
        raise ValueError("fold_at_z: bits must hold 2^m bits")
    if _check_f128(weights, "weights") != 1 << k_skip:
        raise ValueError("fold_at_z: weights must have 2^k_skip entries")
    out = f128_empty(1 << (m - k_skip), bits.device) if out is None else out
    nat.call("flk_zc_fold_at_z", _ptr(bits), m, k_skip, _ptr(weights), _ptr(out), _stream())
    return out


def zerocheck_round_pair(a_mlv, b_mlv, eq_rest, out=None):
    """(G(1), G(inf)) of a multilinear zerocheck round (wire form)."""
    n = _check_f128(a_mlv, "a_mlv")
    if _check_f128(b_mlv, "b_mlv") != n or _check_f128(eq_rest, "eq_rest") != n // 2:
        raise ValueError("zerocheck_round_pair: shapes")
    out = f128_empty(2, a_mlv.device) if out is None else out
    nat.call("flk_zc_round_pair", _ptr(a_mlv), _ptr(b_mlv), n, _ptr(eq_rest), _ptr(out), _stream())
    return out


def zerocheck_round1(a_bits, b_bits, c_bits, m, r_outer):
    """Round-1 univariate-skip message (P_AB, P_C on Lambda, 64 each) for the
    protocol's fixed inner challenges and r_outer (list of m - 13 ints)."""
    for t in (a_bits, b_bits, c_bits):
        if t.numel() != (1 << m) // 8:
            raise ValueError("zerocheck_round1: bit vectors must hold 2^m bits")
    if len(r_outer) != m - 13:
        raise ValueError("zerocheck_round1: r_outer must have m - 13 entries")
    out = f128_empty(128, a_bits.device)
    r = (F128 * max(len(r_outer), 1))(*[F128.from_int(v) for v in r_outer])
    nat.call("flk_zc_round1", _ptr(a_bits), _ptr(b_bits), _ptr(c_bits), m, r, _ptr(out), _ptr(out[64:]), _stream())
    return out[:64], out[64:]


# ---- BLAKE3 R1CS witness (CP/r1cs_hashes/blake3.rs, CP = crates/flock-prover/src/) --
COMPRESSION_WORDS = 28  # cv[8], m[16], counter (lo, hi), block_len, flags


def _witness(blocks, n_blocks_log, lincheck):
    if not (isinstance(blocks, torch.Tensor) and blocks.is_cuda and blocks.dtype == torch.int32
            and blocks.dim() == 2 and blocks.shape[1] == COMPRESSION_WORDS and blocks.is_contiguous()):
        raise ValueError("blocks must be a contiguous CUDA int32 tensor of shape (n, 28)")
    n_blocks = blocks.shape[0]
    if not 3 <= n_blocks_log <= 30:
        raise ValueError("n_blocks_log must be in 3..30 (the lincheck stripe needs >= 8 slots)")
    n_total = 1 << n_blocks_log
    if n_blocks > n_total:
        raise ValueError(f"{n_blocks} blocks > 2^{n_blocks_log} = {n_total} slots")
    z, a, b = (f128_empty(n_total * 128, blocks.device) for _ in range(3))
    zl = torch.empty(n_total * 2048, dtype=torch.uint8, device=blocks.device) if lincheck else None
    nat.call("flk_blake3_witness", _ptr(blocks), n_blocks, n_blocks_log, _ptr(z), _ptr(a), _ptr(b),
             _ptr(zl) if lincheck else None, _stream())
    return z, a, b, zl


def generate_witness_with_ab_packed_and_lincheck(blocks, n_blocks_log):
    """(z, a, b, z_lincheck) for the compressions in `blocks` (one 112-byte record
    per row: cv, m, counter lo/hi, block_len, flags) padded to 2^n_blocks_log
    slots with the all-zero compression (blake3.rs:4617-4647).  z, a, b are F128
    tensors of 2^n_blocks_log * 128 elements (2^14 packed row bits per slot);
    z_lincheck is the uint8 stripe that partial_fold_packed_z reads."""
    return _witness(blocks, n_blocks_log, True)


def generate_witness_with_ab_packed(blocks, n_blocks_log):
    """(z, a, b) as above without the stripe (blake3.rs:2184-2206)."""
    return _witness(blocks, n_blocks_log, False)[:3]


# ---- proof-of-work grinding (CF/challenger.rs, BLAKE3 transcripts) ------------
def _digest_buf(state_digest):
    import ctypes
    d = bytes(state_digest)
    if len(d) != 32:
        raise ValueError("state_digest must be 32 bytes")
    return (ctypes.c_uint8 * 32).from_buffer_copy(d)


def pow_scan(state_digest, start, length, bits):
    """pow_scan: the smallest nonce in [start, start + length) whose BLAKE3 PoW
    hash has `bits` leading zero bits, or None (CF/challenger.rs:1648-1660)."""
    import ctypes
    if not 0 <= bits <= 64:
        raise ValueError("bits must be in 0..64")
    out = ctypes.c_uint64()
    nat.call("flk_pow_scan", _digest_buf(state_digest), start, length, bits, ctypes.byref(out), _stream())
    return None if out.value == 2**64 - 1 else out.value


def grind_pow_nonce(state_digest, bits):
    """The nonce grind_pow finds for this transcript digest (CF/challenger.rs:384-474):
    the smallest satisfying one, 0 when bits == 0.  The caller absorbs it."""
    import ctypes
    if not 0 <= bits <= 48:
        raise ValueError("bits must be in 0..48")
    out = ctypes.c_uint64()
    nat.call("flk_pow_grind", _digest_buf(state_digest), bits, ctypes.byref(out), _stream())
    return out.value


def _u64_host(vals):
    arr = (ctypes_u64 * max(len(vals), 1))(*[int(v) for v in vals])
    return arr


def gather_rows(codeword, row_len, positions, out=None):
    """The opened rows of `positions` (host ints) from a SoA codeword tensor."""
    n_el = _check_f128(codeword, "codeword")
    if row_len <= 0 or n_el % row_len:
        raise ValueError("gather_rows: the codeword length must be a multiple of row_len")
    n = len(positions)
    out = f128_empty(n * row_len, codeword.device) if out is None else out
    nat.call("flk_gather_rows", _ptr(codeword), n_el // row_len, row_len, _u64_host(positions), n, _ptr(out),
             _stream())
    return out


def merkle_multi_proof(tree, num_leaves, positions):
    """merkle_multi_proof (CF/merkle.rs:1020-1060) of a device tree: (n, 32) uint8 tensor of the siblings."""
    import ctypes
    depth = max(num_leaves.bit_length() - 1, 1)
    cap = max(len(positions) * depth, 1)
    out = hashes_empty(cap, tree.device)
    k = ctypes.c_size_t()
    nat.call("flk_merkle_multi_proof", _ptr(tree), num_leaves, _u64_host(positions), len(positions), _ptr(out),
             cap, ctypes.byref(k), _stream())
    return out[: k.value]


class FsChallenger:
    """FsChallenger::with_hash(domain, HashKind::Blake3) (CF/challenger.rs:226-523):
    the Fiat-Shamir transcript on the host; grind_pow runs the nonce search on
    the device (flk_pow_grind), everything else is host BLAKE3."""

    def __init__(self, domain):
        import ctypes
        d = bytes(domain)
        h = ctypes.c_void_p()
        buf = (ctypes.c_uint8 * max(len(d), 1)).from_buffer_copy(d or b"\0")
        nat.call("flk_fs_new", buf, len(d), ctypes.byref(h))
        self._h = h

    @classmethod
    def with_hash(cls, domain):
        return cls(domain)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.load().flk_fs_free(h)
            self._h = None

    def clone(self):
        import ctypes
        c = FsChallenger.__new__(FsChallenger)
        h = ctypes.c_void_p()
        nat.call("flk_fs_clone", self._h, ctypes.byref(h))
        c._h = h
        return c

    @staticmethod
    def _bytes(b):
        import ctypes
        b = bytes(b)
        return (ctypes.c_uint8 * max(len(b), 1)).from_buffer_copy(b or b"\0"), len(b)

    def observe_label(self, label):
        nat.call("flk_fs_observe_label", self._h, *self._bytes(label))

    def observe_bytes(self, data):
        nat.call("flk_fs_observe_bytes", self._h, *self._bytes(data))

    def observe_f128(self, v):
        nat.call("flk_fs_observe_f128", self._h, F128.from_int(v))

    def observe_f128_slice(self, vals):
        arr = (F128 * max(len(vals), 1))(*[F128.from_int(v) for v in vals])
        nat.call("flk_fs_observe_f128_slice", self._h, arr, len(vals))

    def sample_f128(self):
        out = F128()
        import ctypes
        nat.call("flk_fs_sample_f128", self._h, ctypes.byref(out))
        return out.to_int()

    def sample_f128_vec(self, n):
        arr = (F128 * max(n, 1))()
        nat.call("flk_fs_sample_f128_vec", self._h, n, arr)
        return [arr[i].to_int() for i in range(n)]

    def state_digest(self):
        import ctypes
        out = (ctypes.c_uint8 * 32)()
        nat.call("flk_fs_state_digest", self._h, out)
        return bytes(out)

    def sample_distinct_queries(self, block_len, count):
        """sample_distinct_queries (CF/pcs/ligerito.rs:9665-9720), sorted."""
        arr = (ctypes_u64 * max(count, 1))()
        nat.call("flk_fs_sample_distinct_queries", self._h, block_len, count, arr)
        return [int(arr[i]) for i in range(count)]

    def grind_pow(self, bits):
        import ctypes
        n = ctypes.c_uint64()
        nat.call("flk_fs_grind_pow", self._h, bits, ctypes.byref(n), _stream())
        return n.value

    def verify_pow(self, nonce, bits):
        import ctypes
        ok = ctypes.c_int()
        nat.call("flk_fs_verify_pow", self._h, nonce, bits, ctypes.byref(ok))
        return bool(ok.value)


class SumcheckProver:
    """ligerito::SumcheckProver over device-resident f and combined basis
    (CF/pcs/ligerito.rs:8893-9260): new -> first message; fold(r) -> message of
    the folded pair; introduce_new(b_new, h_new) -> message of (f, b_new);
    glue(alpha) -> b += alpha b_new, T += alpha h_new.  Messages are (u_0, u_2)
    device tensors; the running target T is kept as an int."""

    def __init__(self, f, b1, h1):
        if _check_f128(f, "f") != _check_f128(b1, "b1"):
            raise ValueError("SumcheckProver: f and b1 lengths differ")
        self.f = f
        self.combined_basis = b1
        self.t_r = h1
        self.transcript = []
        self.pending_glue = None

    @classmethod
    def new(cls, f, b1, h1):
        inst = cls(f, b1, h1)
        msg = round_msg(inst.f, inst.combined_basis)
        inst.transcript.append(msg)
        return inst, msg

    def fold(self, r):
        if self.pending_glue is not None:
            raise RuntimeError("fold before ordinary glue")
        nf, nb, msg = fold_and_msg(self.f, self.combined_basis, r)
        self.f, self.combined_basis = nf, nb
        self.transcript.append(msg)
        return msg

    def ood(self, challenger, n_samples):
        """The OOD samples of a recursive level (CF/pcs/ligerito.rs:10569-10590):
        per sample z = sample_f128_vec(log n), introduce eq(z) with its
        evaluation, observe (y, u_0, u_2), glue with a sampled beta.  Appends
        the messages to the transcript as (u_0, u_2) int pairs (they are host
        values already) and returns the OOD values."""
        if self.pending_glue is not None:
            raise RuntimeError("ood before ordinary glue")
        n = self.f.shape[0]
        if _check_f128(self.combined_basis, "b") != n or not _pow2(n) or n < 2:
            raise ValueError("ood: f and b must have equal power-of-two length >= 2")
        k = max(n_samples, 1)
        t = F128.from_int(self.t_r)
        ys, ms = (F128 * k)(), (F128 * (2 * k))()
        import ctypes
        nat.call("flk_ligerito_ood", challenger._h, _ptr(self.f), _ptr(self.combined_basis), _log2(n), n_samples,
                 ctypes.byref(t), ys, ms, _stream())
        self.t_r = t.to_int()
        for i in range(n_samples):
            self.transcript.append((ms[2 * i].to_int(), ms[2 * i + 1].to_int()))
        return [ys[i].to_int() for i in range(n_samples)]

    def introduce_new(self, b_new, h_new):
        if _check_f128(b_new, "b_new") != self.f.shape[0]:
            raise ValueError("introduce_new: length mismatch")
        msg = round_msg(self.f, b_new)
        self.transcript.append(msg)
        self.pending_glue = (b_new, h_new)
        return msg

    def glue(self, alpha):
        if self.pending_glue is None:
            raise RuntimeError("glue without introduce_new")
        b_new, h_new = self.pending_glue
        self.pending_glue = None
        glue(self.combined_basis, b_new, alpha)
        self.t_r ^= nat.f128_mul(alpha, h_new)
