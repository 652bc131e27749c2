"""Parity checks of the kernel variants that the library selects by environment
switches (read once per process), run as a child process by
tests/test_gpu_variants.py with one setting each:

  variant_check.py ntt         the commit (codeword + root) at 2^13 x 64 and
                               2^11 x 8 (rate 1/2), the interleaved forward
                               transform and inverse at 2^12 x 64, and a
                               forward/inverse round trip;
  variant_check.py reductions  round_msg (+eval), fold_and_msg, zerocheck round
                               pair, lincheck round message and ring-switch
                               fold_1b_rows at 2^13 and 2^16;
  variant_check.py witness     the BLAKE3 witness (z, a, b, stripe) of 100 and
                               1000 compressions in 2^7 and 2^10 slots;
  variant_check.py round1      zerocheck round 1 at m = 13, 16 and 20.

Uses the C ABI through ctypes only (flk_alloc / flk_copy_*), no torch.  Exit
status 0 means every check matched the oracle; the lines printed name each check.
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


def check_ntt(lib, orc):
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
    return ok


def check_reductions(lib, orc):
    import random
    ok = True

    def run(name, got, want):
        nonlocal ok
        good = same(got, want)
        print(f"{name}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok = ok and good

    for logn in (13, 16):
        n = 1 << logn
        rng = random.Random(logn)
        f, b = fill(10 + logn, n), fill(20 + logn, n)
        df, db, dout = Dev(lib, f.nbytes), Dev(lib, b.nbytes), Dev(lib, 128 * 16)
        df.put(f)
        db.put(b)
        nat.check(lib.flk_sc_round_msg(df.p, db.p, n, dout.p, None), "round_msg")
        run(f"round_msg 2^{logn}", dout.get(F128_DT, 2), orc.round_msg(f, b))
        nat.check(lib.flk_sc_round_msg_and_eval(df.p, db.p, n, dout.p, None), "round_msg_and_eval")
        run(f"round_msg_and_eval 2^{logn}", dout.get(F128_DT, 3), orc.round_msg_and_eval(f, b))
        r = rng.getrandbits(128)
        dfo, dbo = Dev(lib, f.nbytes // 2), Dev(lib, b.nbytes // 2)
        nat.check(lib.flk_sc_fold_and_msg(df.p, db.p, n, nat.F128.from_int(r), dfo.p, dbo.p, dout.p, None),
                  "fold_and_msg")
        wfo, wbo, wout = orc.fold_and_msg(f, b, r)
        run(f"fold_and_msg 2^{logn} (folds)",
            np.concatenate([dfo.get(F128_DT, n // 2), dbo.get(F128_DT, n // 2)]), np.concatenate([wfo, wbo]))
        run(f"fold_and_msg 2^{logn} (message)", dout.get(F128_DT, 2), wout)
        eq_rest = orc.build_eq([rng.getrandbits(128) for _ in range(logn - 1)])
        de = Dev(lib, eq_rest.nbytes)
        de.put(eq_rest)
        nat.check(lib.flk_zc_round_pair(df.p, db.p, n, de.p, dout.p, None), "zc_round_pair")
        run(f"zc_round_pair 2^{logn}", dout.get(F128_DT, 2), orc.zc_round_pair(f, b, eq_rest))
        nat.check(lib.flk_lc_round_msg_msb(df.p, db.p, n, dout.p, None), "lc_round_msg_msb")
        run(f"lc_round_msg_msb 2^{logn}", dout.get(F128_DT, 2), orc.lc_round_msg_msb(f, b))
        nat.check(lib.flk_rs_fold_1b_rows(df.p, db.p, n, dout.p, None), "rs_fold_1b_rows")
        run(f"rs_fold_1b_rows 2^{logn}", dout.get(F128_DT, 128), orc.rs_fold_1b_rows(f, b))
        for d in (df, db, dout, dfo, dbo, de):
            d.free()
    return ok


def check_witness(lib, orc):
    """flk_blake3_witness (z, a, b, stripe) at 2^7 and 2^10 slots, ragged."""
    from oracle_ffi import COMPRESSION_DT
    ok = True
    for n_log, n in ((7, 100), (10, 1000)):
        n_total = 1 << n_log
        c = np.zeros(n, COMPRESSION_DT)
        c.view(np.uint32).reshape(n, 28)[:] = np.random.default_rng(n).integers(0, 2**32, (n, 28), dtype=np.uint32)
        want = orc.witness(c, n_log)
        dc = Dev(lib, c.nbytes)
        dc.put(c)
        outs = [Dev(lib, n_total * 2048) for _ in range(4)]
        nat.check(lib.flk_blake3_witness(dc.p, n, n_log, outs[0].p, outs[1].p, outs[2].p, outs[3].p, None),
                  "blake3_witness")
        for name, d, w in zip(("z", "a", "b"), outs[:3], want[:3]):
            good = bool(np.array_equal(d.get(np.uint64, n_total * 256), w))
            print(f"witness {name} 2^{n_log}: {'ok' if good else 'MISMATCH'}", flush=True)
            ok = ok and good
        good = bool(np.array_equal(outs[3].get(np.uint8, n_total * 2048), want[3]))
        print(f"witness stripe 2^{n_log}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok = ok and good
        for d in [dc] + outs:
            d.free()
    return ok


def check_round1(lib, orc):
    """flk_zc_round1 at m = 13, 16, 20 (the protocol's fixed inner challenges) against
    round1_naive, plus the all-ones rows of an honest-looking witness (a AND b = c)."""
    import random
    import flock_verifier as fv
    ok = True
    for m in (13, 16, 20):
        rng = random.Random(m)
        nb = (1 << m) // 8
        a, b, c = (np.frombuffer(rng.randbytes(nb), np.uint8).copy() for _ in range(3))
        if m == 16:
            a[: nb // 4] = 0xFF
            c = a & b
        r_outer = [rng.getrandbits(128) for _ in range(m - 13)]
        r = [rng.getrandbits(128) for _ in range(6)] + fv.fixed_challenges(orc) + r_outer
        want_ab, want_c = orc.zc_round1(a, b, c, m, 6, r)
        da, db, dc, dout = Dev(lib, nb), Dev(lib, nb), Dev(lib, nb), Dev(lib, 128 * 16)
        da.put(a)
        db.put(b)
        dc.put(c)
        ro = (nat.F128 * max(len(r_outer), 1))(*[nat.F128.from_int(v) for v in r_outer])
        nat.check(lib.flk_zc_round1(da.p, db.p, dc.p, m, ro, dout.p, ctypes.c_void_p(dout.p.value + 64 * 16), None),
                  "zc_round1")
        got = dout.get(F128_DT, 128)
        good = same(got[:64], want_ab) and same(got[64:], want_c)
        print(f"zc_round1 m={m}: {'ok' if good else 'MISMATCH'}", flush=True)
        ok = ok and good
        for d in (da, db, dc, dout):
            d.free()
    return ok


def main():
    lib = nat.load()
    orc = Oracle()
    mode = sys.argv[1] if len(sys.argv) > 1 else "ntt"
    check = {"ntt": check_ntt, "reductions": check_reductions, "witness": check_witness,
             "round1": check_round1}[mode]
    ok = check(lib, orc)
    return 0 if ok else 1


if __name__ == "__main__":
    sys.exit(main())
