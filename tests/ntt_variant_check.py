"""Parity check of the NTT product-kernel variants that the library selects by
environment switches (read once per process), run as a child process by
tests/test_gpu_ntt_variants.py with one setting each.

Checks, against the oracle: the commit (codeword + root) at 2^13 x 64 and
2^11 x 8 (rate 1/2), the interleaved forward transform and inverse at
2^12 x 64, and a forward/inverse round trip.  Uses the C ABI through ctypes
only (flk_alloc / flk_copy_*), no torch.  Exit status 0 means every check
matched; the lines printed name each check.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

import flock_challenge_multi_amd._native as nat  # noqa: E402
from oracle_ffi import F128_DT, Oracle  # noqa: E402


class Dev:
    def __init__(self, lib, nbytes):
        self.lib = lib
        self.p = ctypes.c_void_p()
        nat.check(lib.flk_alloc(nbytes, ctypes.byref(self.p)), "flk_alloc")
        self.nbytes = nbytes

    def put(self, a):
        a = np.ascontiguousarray(a)
        nat.check(self.lib.flk_copy_to_device(self.p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, None),
                  "flk_copy_to_device")

    def get(self, dtype, n):
        out = np.zeros(n, dtype=dtype)
        nat.check(self.lib.flk_copy_to_host(out.ctypes.data_as(ctypes.c_void_p), self.p, out.nbytes, None),
                  "flk_copy_to_host")
        nat.check(self.lib.flk_synchronize(None), "flk_synchronize")
        return out

    def free(self):
        self.lib.flk_free(self.p)


def fill(seed, n):
    rng = np.random.default_rng(seed)
    a = np.zeros(n, dtype=F128_DT)
    a["lo"] = rng.integers(0, 2**64, n, dtype=np.uint64)
    a["hi"] = rng.integers(0, 2**64, n, dtype=np.uint64)
    return a


def same(a, b):
    return a.shape == b.shape and bool((a["lo"] == b["lo"]).all() and (a["hi"] == b["hi"]).all())


def main():
    lib = nat.load()
    orc = Oracle()
    ok = True
    for log_cols, log_lanes in [(13, 6), (11, 3)]:
        lanes = 1 << log_lanes
        msg = fill(1000 + log_cols, lanes << log_cols)
        n_cw = len(msg) * 2
        n_pos = n_cw // lanes
        want_cw, _, want_root = orc.commit(msg, lanes, 1)
        dm, dc, dt = Dev(lib, msg.nbytes), Dev(lib, n_cw * 16), Dev(lib, (2 * n_pos - 1) * 32)
        dm.put(msg)
        root = (ctypes.c_uint8 * 32)()
        nat.check(lib.flk_commit(dm.p, log_cols, log_lanes, 1, nat.FLK_HASH_BLAKE3, dc.p, dt.p, root, None),
                  "flk_commit")
        good = same(dc.get(F128_DT, n_cw), want_cw) and bytes(root) == want_root
        print(f"commit 2^{log_cols} x {lanes}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok = ok and good
        for d in (dm, dc, dt):
            d.free()
    log_d, lanes = 12, 64
    data = fill(77, (1 << log_d) * lanes)
    want_fwd = orc.forward_interleaved(log_d, data, lanes, 0)
    cols = data.reshape(1 << log_d, lanes)
    want_inv = np.stack([orc.inverse(log_d, np.ascontiguousarray(cols[:, c])) for c in range(lanes)],
                        axis=1).reshape(-1)
    dd = Dev(lib, data.nbytes)
    dd.put(data)
    nat.check(lib.flk_ntt_forward_interleaved(log_d, dd.p, len(data), lanes, 0, None), "forward")
    fwd = dd.get(F128_DT, len(data))
    good = same(fwd, want_fwd)
    print(f"forward 2^{log_d} x {lanes}: {'ok' if good else 'MISMATCH'}", flush=True)
    ok = ok and good
    nat.check(lib.flk_ntt_inverse_interleaved(log_d, dd.p, len(data), lanes, None), "inverse")
    good = same(dd.get(F128_DT, len(data)), data)
    print(f"forward then inverse 2^{log_d} x {lanes}: {'ok' if good else 'MISMATCH'}", flush=True)
    ok = ok and good
    dd.put(data)
    nat.check(lib.flk_ntt_inverse_interleaved(log_d, dd.p, len(data), lanes, None), "inverse")
    good = same(dd.get(F128_DT, len(data)), want_inv)
    print(f"inverse 2^{log_d} x {lanes}: {'ok' if good else 'MISMATCH'}", flush=True)
    ok = ok and good
    dd.free()
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
