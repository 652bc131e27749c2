"""ctypes binding of the gfx950 device library `libflock_gpu.so` (C ABI in
include/flock_gpu.h).

The library is built in-tree by `flock_challenge_multi_amd/build.py`.  Loading
fails loudly when it is missing: there is no CPU fallback in the product path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libflock_gpu.so")

FLK_OK = 0
FLK_E_SHAPE = 1
FLK_E_DEVICE = 2
FLK_E_DISABLED = 3
FLK_E_NOMEM = 4
FLK_HASH_SHA256 = 0
FLK_HASH_BLAKE3 = 1


class F128(ctypes.Structure):
    """flk_f128: `lo` (x^0..x^63) then `hi` (x^64..x^127)."""

    _fields_ = [("lo", ctypes.c_uint64), ("hi", ctypes.c_uint64)]

    @classmethod
    def from_int(cls, v):
        return cls(v & (2**64 - 1), (v >> 64) & (2**64 - 1))

    def to_int(self):
        return self.lo | (self.hi << 64)


class FlockGpuError(RuntimeError):
    def __init__(self, code, func, msg):
        super().__init__(f"{func} failed with status {code}: {msg}")
        self.code = code


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_I = ctypes.c_int
_F = F128

# name -> (restype, argtypes)
SIGNATURES = {
    "flk_device_available": (_I, []),
    "flk_is_disabled": (_I, []),
    "flk_undisable": (None, []),
    "flk_set_enabled": (None, [_I]),
    "flk_last_error": (ctypes.c_char_p, []),
    "flk_version": (ctypes.c_char_p, []),
    "flk_f128_mul": (_F, [_F, _F]),
    "flk_alloc": (_I, [_SZ, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_free": (_I, [_P]),
    "flk_alloc_async": (_I, [_SZ, _P, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_free_async": (_I, [_P, _P]),
    "flk_copy_to_device": (_I, [_P, _P, _SZ, _P]),
    "flk_copy_to_host": (_I, [_P, _P, _SZ, _P]),
    "flk_synchronize": (_I, [_P]),
    "flk_ntt_twiddles": (_I, [_SZ, _P]),
    "flk_ntt_set_schedule": (_I, [_SZ, _SZ, _P, _SZ]),
    "flk_ntt_get_schedule": (_I, [_SZ, _SZ, _P, _SZ, _P]),
    "flk_ntt_forward_interleaved": (_I, [_SZ, _P, _SZ, _SZ, _SZ, _P]),
    "flk_ntt_rs_encode_interleaved": (_I, [_SZ, _P, _SZ, _P, _SZ, _SZ, _P]),
    "flk_ntt_inverse_interleaved": (_I, [_SZ, _P, _SZ, _SZ, _P]),
    "flk_merkle_hash_leaves": (_I, [_P, _SZ, _SZ, _P, _P]),
    "flk_merkle_hash_pairs_level": (_I, [_P, _SZ, _P, _P]),
    "flk_merkle_tree": (_I, [_P, _SZ, _SZ, _I, _P, _P]),
    "flk_merkle_begin": (_I, [_P, _SZ, _SZ, _P, _SZ, _SZ, _P, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_merkle_session_leaves": (_SZ, [_P]),
    "flk_merkle_commit_leaves": (_I, [_P, _SZ, _SZ]),
    "flk_merkle_commit_parent_levels": (_I, [_P]),
    "flk_merkle_finish": (_I, [_P, ctypes.POINTER(ctypes.c_double)]),
    "flk_merkle_discard": (None, [_P]),
    "flk_prewire": (_I, [_P, _SZ]),
    "flk_commit": (_I, [_P, _SZ, _SZ, _SZ, _I, _P, _P, _P, _P]),
    "flk_build_eq": (_I, [_P, _SZ, _P, _P]),
    "flk_sc_round_msg": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_sc_round_msg_and_eval": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_sc_round_msg_noalloc": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_sc_fold_and_msg": (_I, [_P, _P, _SZ, _F, _P, _P, _P, _P]),
    "flk_sc_glue": (_I, [_P, _P, _SZ, _F, _P]),
    "flk_partial_eval_lsb": (_I, [_P, _SZ, _P, _SZ, _P, _P]),
    "flk_induce_sumcheck_poly": (_I, [_SZ, _P, _SZ, _P, _SZ, _P, _P, _SZ, _P, _P, _P]),
    "flk_lc_partial_fold": (_I, [_P, _SZ, _SZ, _P, _P, _P]),
    "flk_lc_round_msg_msb": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_lc_bind_top": (_I, [_P, _SZ, _F, _P]),
    "flk_rs_fold_1b_rows": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_rs_fold_b128_elems": (_I, [_P, _P, _SZ, _P, _P]),
    "flk_rs_prove_batched": (_I, [_P, _P, _SZ, _P, _SZ, _P, _P, _P, _P, _P]),
    "flk_ligerito_ood": (_I, [_P, _P, _P, _SZ, _SZ, _P, _P, _P, _P]),
    "flk_pcs_combined_basis_and_target": (_I, [_P, _P, _SZ, _P, _SZ, _P, _P, _P, _P, _P, _P]),
    "flk_zc_lagrange_weights": (_I, [_SZ, _F, _P]),
    "flk_zc_fold_at_z": (_I, [_P, _SZ, _SZ, _P, _P, _P]),
    "flk_zc_round_pair": (_I, [_P, _P, _SZ, _P, _P, _P]),
    "flk_zc_round1": (_I, [_P, _P, _P, _SZ, _P, _P, _P, _P]),
    "flk_blake3_witness": (_I, [_P, _SZ, _SZ, _P, _P, _P, _P, _P]),
    "flk_pow_scan": (_I, [_P, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), _P]),
    "flk_pow_grind": (_I, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), _P]),
    "flk_fs_new": (_I, [_P, _SZ, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_fs_clone": (_I, [_P, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_fs_free": (None, [_P]),
    "flk_fs_observe_label": (_I, [_P, _P, _SZ]),
    "flk_fs_observe_f128": (_I, [_P, _F]),
    "flk_fs_observe_f128_slice": (_I, [_P, _P, _SZ]),
    "flk_fs_observe_bytes": (_I, [_P, _P, _SZ]),
    "flk_fs_sample_f128": (_I, [_P, ctypes.POINTER(_F)]),
    "flk_fs_sample_f128_vec": (_I, [_P, _SZ, _P]),
    "flk_fs_state_digest": (_I, [_P, _P]),
    "flk_fs_grind_pow": (_I, [_P, ctypes.c_uint32, ctypes.POINTER(ctypes.c_uint64), _P]),
    "flk_fs_verify_pow": (_I, [_P, ctypes.c_uint64, ctypes.c_uint32, ctypes.POINTER(ctypes.c_int)]),
    "flk_fs_sample_distinct_queries": (_I, [_P, ctypes.c_uint64, _SZ, _P]),
    "flk_gather_rows": (_I, [_P, _SZ, _SZ, _P, _SZ, _P, _P]),
    "flk_merkle_multi_proof": (_I, [_P, _SZ, _P, _SZ, _P, _SZ, ctypes.POINTER(_SZ), _P]),
    "flk_copy_device_to_device": (_I, [_P, _P, _SZ, _P]),
    "flk_blake3_r1cs_new": (_I, [_SZ, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_r1cs_free": (None, [_P]),
    "flk_r1cs_shape": (_I, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ),
                            ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "flk_r1cs_statement_digest": (_I, [_P, _P]),
    "flk_blake3_statement_digest": (_I, [_SZ, _P]),
    "flk_lincheck_comb": (_I, [_P, _F, _P, _SZ, _F, _F, _P, _P]),
    "flk_zerocheck_prove": (_I, [_P, _P, _P, _P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P]),
    "flk_lincheck_prove": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "flk_blake3_setup_new": (_I, [_SZ, _I, ctypes.POINTER(ctypes.c_void_p)]),
    "flk_blake3_setup_free": (None, [_P]),
    "flk_blake3_setup_info": (_I, [_P, ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), ctypes.POINTER(_SZ), _P]),
    "flk_blake3_proof_bound": (_I, [_P, ctypes.POINTER(_SZ)]),
    "flk_inject_device_failures": (None, [ctypes.c_uint, ctypes.c_uint]),
    "flk_read_to_host": (_I, [_P, _P, _SZ, _P]),
    "flk_read_arm": (_I, [_P, _SZ]),
    "flk_read_armed": (_I, [_P, _SZ, _P]),
    "flk_zc_fold_round": (_I, [_P, _P, _SZ, _P, _P, _P, _SZ, _P, _P, _P, _P, _P]),
    "flk_zc_fold_at_z_round": (_I, [_P, _P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P]),
    "flk_zc_fold_at_z_bind_round": (_I, [_P, _P, _SZ, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    "flk_blake3_setup_arena_stats": (_I, [_P] + [ctypes.POINTER(_SZ)] * 4),
    "flk_prover_pool_new": (_I, [ctypes.c_int, _SZ, _SZ, ctypes.POINTER(_P)]),
    "flk_prover_pool_free": (None, [_P]),
    "flk_prover_pool_prepare": (_I, [_P, _SZ]),
    "flk_prover_pool_submit": (_I, [_P, _P, _SZ, _P, _SZ, ctypes.POINTER(ctypes.c_uint64)]),
    "flk_prover_pool_wait": (_I, [_P, ctypes.c_uint64, _P, _SZ, ctypes.POINTER(_SZ), _P]),
    "flk_prover_pool_stats": (_I, [_P] + [ctypes.POINTER(_SZ)] * 4),
    "flk_blake3_prove": (_I, [_P, _P, _SZ, _P, _SZ, _P, _SZ, ctypes.POINTER(_SZ), _P, _P]),
}

_lib = None


def load():
    """Load libflock_gpu.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python flock_challenge_multi_amd/build.py` "
            "(the device path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status, func):
    if status != FLK_OK:
        msg = load().flk_last_error()
        raise FlockGpuError(status, func, msg.decode() if msg else "")
    return status


def f128_mul(a, b):
    """Host scalar GF(2^128) product of two ints (device library's host path)."""
    return load().flk_f128_mul(F128.from_int(a), F128.from_int(b)).to_int()


def call(name, *args):
    """Call an int-returning entry point and raise on a non-OK status."""
    return check(getattr(load(), name)(*args), name)
