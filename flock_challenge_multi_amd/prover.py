"""Python mirror of the reference's BLAKE3 proving interface over the device
library (CP = crates/flock-prover/src/, CF = crates/flock-core/src/):

  Blake3Setup(n_blocks, profile)     Blake3Setup::new / with_profile
                                     (CP/r1cs_hashes/blake3.rs:4706-4788)
  Blake3Setup.prove(blocks, domain)  the proof bytes of prove_fast / prove_ligerito
                                     (CP/prover.rs:373-472) serialised by
                                     R1csProofBundleLigerito::to_bytes
                                     (CP/proof_io.rs:160-198)
  R1cs.blake3(n_blocks_log)          build_block_r1cs (blake3.rs:1531-1545), its
                                     statement_digest (CF/r1cs.rs:334-385) and the
                                     lincheck comb vector (CF/lincheck.rs:5590-5660)
  zerocheck_prove / lincheck_prove   zerocheck::prove_packed_padded (CF/zerocheck.rs:384-535),
                                     lincheck::prove_padded (CF/lincheck.rs:5345-5878)

Inputs are torch CUDA tensors; work runs on torch's current stream.  Shape
errors raise ValueError, device failures FlockGpuError.
"""
import ctypes

import numpy as np
import torch

from . import _native as nat
from . import hotpath as hp
from ._native import F128

DOMAIN = b"flock-bench-v0"  # the benchmark's Fiat-Shamir domain (benchmark-tools/common/src/lib.rs:2)
PROFILES = {"fast": 0, "slim": 1, "secure": 2}
STAGES = ("witness+commit", "zerocheck", "lincheck", "ring_switch+combine", "ligerito_open", "serialize")


def _ints(arr, n):
    return [arr[i].to_int() for i in range(n)]


class R1cs:
    """The BLAKE3 R1CS of 2^n_blocks_log compressions with its device circuit."""

    def __init__(self, n_blocks_log):
        h = ctypes.c_void_p()
        nat.call("flk_blake3_r1cs_new", n_blocks_log, ctypes.byref(h))
        self._h = h

    @classmethod
    def blake3(cls, n_blocks_log):
        return cls(n_blocks_log)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.load().flk_r1cs_free(h)
            self._h = None

    def shape(self):
        v = [ctypes.c_size_t() for _ in range(5)]
        nat.call("flk_r1cs_shape", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("m", "k_log", "k_skip", "nnz_a", "nnz_b"), (x.value for x in v)))

    def statement_digest(self):
        out = (ctypes.c_uint8 * 32)()
        nat.call("flk_r1cs_statement_digest", self._h, out)
        return bytes(out)

    def lincheck_comb(self, z_skip, x_inner_rest, alpha, beta, out=None):
        """comb = fold_alpha_batched(alpha, quirky eq(z_skip, x_inner_rest)) + beta at
        the constant column, a (2^14, 2) int64 device tensor."""
        if out is None:
            out = hp.f128_empty(1 << 14)
        xr = (F128 * len(x_inner_rest))(*[F128.from_int(v) for v in x_inner_rest])
        nat.call("flk_lincheck_comb", self._h, F128.from_int(z_skip), xr, len(x_inner_rest),
                 F128.from_int(alpha), F128.from_int(beta), hp._ptr(out), hp._stream())
        return out


def _bits_ptr(t, m, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous CUDA tensor")
    if t.numel() * t.element_size() != (1 << m) // 8:
        raise ValueError(f"{name} must hold 2^m bits")
    return hp._ptr(t)


def zerocheck_prove(challenger, a_bits, b_bits, c_bits, m):
    """zerocheck::prove_packed_padded on the challenger's transcript.  Returns
    (proof, claim) dicts of ints: proof = round1_ab, round1_c, multilinear_rounds,
    final_a_eval, final_b_eval, final_c_eval; claim = z, mlv_challenges, r_rest."""
    if not 13 <= m <= 40:
        raise ValueError("zerocheck_prove: 13 <= m <= 40")
    n = m - 6
    r1ab, r1c = (F128 * 64)(), (F128 * 64)()
    rounds, finals = (F128 * (2 * n))(), (F128 * 3)()
    z, mlv, rr = F128(), (F128 * n)(), (F128 * n)()
    nat.call("flk_zerocheck_prove", challenger._h, _bits_ptr(a_bits, m, "a"), _bits_ptr(b_bits, m, "b"),
             _bits_ptr(c_bits, m, "c"), m, r1ab, r1c, rounds, finals, ctypes.byref(z), mlv, rr, hp._stream())
    rs = _ints(rounds, 2 * n)
    fin = _ints(finals, 3)
    proof = {"round1_ab": _ints(r1ab, 64), "round1_c": _ints(r1c, 64),
             "multilinear_rounds": [(rs[2 * i], rs[2 * i + 1]) for i in range(n)],
             "final_a_eval": fin[0], "final_b_eval": fin[1], "final_c_eval": fin[2]}
    claim = {"z": z.to_int(), "mlv_challenges": _ints(mlv, n), "r_rest": _ints(rr, n)}
    return proof, claim


def lincheck_prove(challenger, r1cs, z_stripe, x_ab):
    """lincheck::prove_padded with the R1cs circuit; x_ab = [z_skip] + x_inner_rest (8)
    + x_outer (m - 14).  Returns (proof, claim): rounds, z_partial; r_inner_skip,
    r_inner_rest, w."""
    m = r1cs.shape()["m"]
    if len(x_ab) != 1 + 8 + (m - 14):
        raise ValueError("lincheck_prove: x_ab needs 1 + 8 + (m - 14) coordinates")
    if not (isinstance(z_stripe, torch.Tensor) and z_stripe.is_cuda and z_stripe.numel() == (1 << m) // 8):
        raise ValueError("lincheck_prove: z_stripe must be the 2^m-bit device stripe")
    xs = (F128 * len(x_ab))(*[F128.from_int(v) for v in x_ab])
    rounds, zp, rs, rr, w = (F128 * 16)(), (F128 * 64)(), F128(), (F128 * 8)(), F128()
    nat.call("flk_lincheck_prove", challenger._h, r1cs._h, hp._ptr(z_stripe), xs, rounds, zp, ctypes.byref(rs),
             rr, ctypes.byref(w), hp._stream())
    q = _ints(rounds, 16)
    proof = {"rounds": [(q[2 * i], q[2 * i + 1]) for i in range(8)], "z_partial": _ints(zp, 64)}
    claim = {"r_inner_skip": rs.to_int(), "r_inner_rest": _ints(rr, 8), "w": w.to_int()}
    return proof, claim


class Blake3Setup:
    """Blake3Setup for n_blocks compressions (m = 14 + ceil(log2(max(n_blocks, 8))))."""

    def __init__(self, n_blocks, profile="fast"):
        h = ctypes.c_void_p()
        nat.call("flk_blake3_setup_new", n_blocks, PROFILES[profile], ctypes.byref(h))
        self._h = h
        m, nbl, dev = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        dig = (ctypes.c_uint8 * 32)()
        nat.call("flk_blake3_setup_info", h, ctypes.byref(m), ctypes.byref(nbl), ctypes.byref(dev), dig)
        self.m, self.n_blocks_log, self.device_bytes = m.value, nbl.value, dev.value
        self.statement_digest = bytes(dig)
        self.profile = profile
        bound = ctypes.c_size_t()
        nat.call("flk_blake3_proof_bound", h, ctypes.byref(bound))
        self._cap = bound.value

    def arena_stats(self):
        """{capacity, peak, n_arenas, fallbacks} of the setup's device arenas (bytes)."""
        v = [ctypes.c_size_t() for _ in range(4)]
        nat.call("flk_blake3_setup_arena_stats", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("capacity", "peak", "n_arenas", "fallbacks"), (x.value for x in v)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.load().flk_blake3_setup_free(h)
            self._h = None

    def prove(self, blocks, domain=DOMAIN, cap=None, stream=None):
        """Proof bytes for the device compressions `blocks` (a contiguous CUDA int32
        or uint32 tensor of shape (n, 28), n <= 2^n_blocks_log), and the per-stage
        wall times in seconds (STAGES order)."""
        if not (isinstance(blocks, torch.Tensor) and blocks.is_cuda and blocks.dim() == 2
                and blocks.shape[1] == 28 and blocks.element_size() == 4 and blocks.is_contiguous()):
            raise ValueError("blocks must be a contiguous CUDA 32-bit tensor of shape (n, 28)")
        if blocks.shape[0] > (1 << self.n_blocks_log):
            raise ValueError("more compressions than the setup's slots")
        cap = cap or self._cap
        plen = ctypes.c_size_t()
        stages = (ctypes.c_double * 6)()
        d = bytes(domain)
        dbuf = (ctypes.c_uint8 * max(len(d), 1)).from_buffer_copy(d or b"\0")
        st = ctypes.c_void_p(stream.cuda_stream) if stream is not None else hp._stream()
        buf = np.empty(cap, np.uint8)  # not zero-filled: the library writes plen bytes
        nat.call("flk_blake3_prove", self._h, hp._ptr(blocks), blocks.shape[0], dbuf, len(d),
                 buf.ctypes.data_as(ctypes.c_void_p), cap, ctypes.byref(plen), stages, st)
        return buf[:plen.value].tobytes(), list(stages)


class ProverPool:
    """Batched and served proving (flk_prover_pool_*): `workers` host threads, each on
    its own stream, proving batches of any supported size through one device memory
    budget (bytes; None = 60% of the free device memory), admitted in submission
    order.  submit() returns a ticket at once; wait(ticket) returns the proof bytes
    and (seconds queued, seconds proving, seconds submit to done)."""

    def __init__(self, profile="fast", workers=2, hbm_budget=None):
        h = ctypes.c_void_p()
        nat.call("flk_prover_pool_new", PROFILES[profile], workers, hbm_budget or 0, ctypes.byref(h))
        self._h = h
        self._held = {}  # ticket -> the input tensor, kept alive until the job is waited for
        self.workers = workers

    def prepare(self, n_blocks):
        nat.call("flk_prover_pool_prepare", self._h, n_blocks)

    def submit(self, blocks, domain=DOMAIN):
        if not (isinstance(blocks, torch.Tensor) and blocks.is_cuda and blocks.dim() == 2
                and blocks.shape[1] == 28 and blocks.element_size() == 4 and blocks.is_contiguous()):
            raise ValueError("blocks must be a contiguous CUDA 32-bit tensor of shape (n, 28)")
        d = bytes(domain)
        dbuf = (ctypes.c_uint8 * max(len(d), 1)).from_buffer_copy(d or b"\0")
        t = ctypes.c_uint64()
        nat.call("flk_prover_pool_submit", self._h, hp._ptr(blocks), blocks.shape[0], dbuf, len(d), ctypes.byref(t))
        self._held[t.value] = blocks
        return t.value

    def wait(self, ticket):
        n = ctypes.c_size_t()
        times = (ctypes.c_double * 3)()
        buf = np.empty(1 << 20, np.uint8)
        rc = nat.load().flk_prover_pool_wait(self._h, ticket, buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                             ctypes.byref(n), times)
        if rc == nat.FLK_E_SHAPE and n.value > buf.size:  # a proof larger than 1 MiB: fetch it again
            buf = np.empty(n.value, np.uint8)
            rc = nat.load().flk_prover_pool_wait(self._h, ticket, buf.ctypes.data_as(ctypes.c_void_p), buf.size,
                                                 ctypes.byref(n), times)
        self._held.pop(ticket, None)
        nat.check(rc, "flk_prover_pool_wait")
        return buf[:n.value].tobytes(), tuple(times)

    def stats(self):
        v = [ctypes.c_size_t() for _ in range(4)]
        nat.call("flk_prover_pool_stats", self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("budget", "peak_reserved", "peak_running", "completed"), (x.value for x in v)))

    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.load().flk_prover_pool_free(h)
            self._h = None

    def __del__(self):
        self.close()
