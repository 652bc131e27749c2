"""GPU parity: eq tables, LSB folds and the Ligerito sumcheck reductions on
gfx950 against the oracle, bit-exact, through the C ABI."""
import random
import sys

import numpy as np
import pytest
import torch

import pyfield as pf
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import from_dev, same, to_dev
from oracle_ffi import f128_to_ints

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("d", [0, 1, 3, 8, 9, 12, 17, 20])
def test_build_eq_matches_oracle(orc, dev, d):
    _log(f"build_eq d={d}")
    rng = random.Random(d)
    pt = [rng.getrandbits(128) for _ in range(d)]
    got = hp.build_eq_table(pt)
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.build_eq(pt))


@pytest.mark.parametrize("logn", [1, 2, 5, 10, 16, 21])
def test_round_msg_matches_oracle(orc, dev, logn):
    _log(f"round_msg logn={logn}")
    f, b = orc.fill(100 + logn, 1 << logn), orc.fill(200 + logn, 1 << logn)
    got = hp.round_msg(to_dev(f), to_dev(b))
    got3 = hp.round_msg_and_eval(to_dev(f), to_dev(b))
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.round_msg(f, b))
    assert same(from_dev(got3), orc.round_msg_and_eval(f, b))


@pytest.mark.parametrize("logn", [2, 3, 6, 12, 18, 22])
def test_fold_and_msg_matches_oracle(orc, dev, logn):
    _log(f"fold_and_msg logn={logn}")
    rng = random.Random(logn)
    r = rng.getrandbits(128)
    f, b = orc.fill(300 + logn, 1 << logn), orc.fill(400 + logn, 1 << logn)
    fo, bo, msg = orc.fold_and_msg(f, b, r)
    gf, gb, gm = hp.fold_and_msg(to_dev(f), to_dev(b), r)
    torch.cuda.synchronize()
    assert same(from_dev(gf), fo) and same(from_dev(gb), bo) and same(from_dev(gm), msg)


@pytest.mark.parametrize("n", [1, 7, 256, 100003])
def test_glue_matches_oracle(orc, dev, n):
    b, bn = orc.fill(5 + n, n), orc.fill(6 + n, n)
    alpha = random.Random(n).getrandbits(128)
    bt = to_dev(b)
    hp.glue(bt, to_dev(bn), alpha)
    torch.cuda.synchronize()
    assert same(from_dev(bt), orc.glue(b, bn, alpha))


@pytest.mark.parametrize("logn,k", [(1, 1), (4, 2), (8, 8), (12, 5), (16, 9), (20, 13)])
def test_partial_eval_matches_oracle(orc, dev, logn, k):
    _log(f"partial_eval logn={logn} k={k}")
    rng = random.Random(logn * 31 + k)
    ev = orc.fill(logn + k, 1 << logn)
    rs = [rng.getrandbits(128) for _ in range(k)]
    got = hp.partial_eval_lsb(to_dev(ev), rs)
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.partial_eval_lsb(ev, rs))


def test_sumcheck_prover_flow_matches_oracle(orc, dev):
    """new -> fold x3 -> introduce_new -> glue -> fold, as the Ligerito prover
    sequences it (CF/pcs/ligerito.rs:8906-9260), against the oracle step by step."""
    rng = random.Random(9)
    n = 1 << 14
    f, b1 = orc.fill(1, n), orc.fill(2, n)
    h1 = 0
    for x, y in zip(f128_to_ints(f), f128_to_ints(b1)):
        h1 ^= pf.mul(x, y)
    prover, msg = hp.SumcheckProver.new(to_dev(f), to_dev(b1), h1)
    assert same(from_dev(msg), orc.round_msg(f, b1))
    cf, cb = f, b1
    for _ in range(3):
        r = rng.getrandbits(128)
        msg = prover.fold(r)
        cf, cb, want = orc.fold_and_msg(cf, cb, r)
        assert same(from_dev(msg), want)
    b_new = orc.fill(3, len(cf))
    h_new = rng.getrandbits(128)
    msg = prover.introduce_new(to_dev(b_new), h_new)
    assert same(from_dev(msg), orc.round_msg(cf, b_new))
    alpha = rng.getrandbits(128)
    prover.glue(alpha)
    cb = orc.glue(cb, b_new, alpha)
    assert same(from_dev(prover.combined_basis), cb)
    assert prover.t_r == h1 ^ pf.mul(alpha, h_new)
    r = rng.getrandbits(128)
    msg = prover.fold(r)
    cf, cb, want = orc.fold_and_msg(cf, cb, r)
    torch.cuda.synchronize()
    assert same(from_dev(msg), want)
    assert same(from_dev(prover.f), cf) and same(from_dev(prover.combined_basis), cb)


@pytest.mark.parametrize("log_cols,n_v,nq,n_alpha", [(0, 3, 1, 1), (4, 3, 5, 3), (9, 3, 40, 6), (13, 6, 71, 7),
                                                   (19, 6, 218, 8)])
def test_induce_sumcheck_poly_matches_oracle(orc, dev, log_cols, n_v, nq, n_alpha):
    _log(f"induce cols=2^{log_cols} n_v={n_v} nq={nq}")
    rng = random.Random(log_cols * 100 + nq)
    rows = orc.fill(7 + log_cols, nq << n_v).reshape(nq, 1 << n_v)
    v = [rng.getrandbits(128) for _ in range(n_v)]
    alpha = [rng.getrandbits(128) for _ in range(n_alpha)]
    queries = [rng.randrange(1 << (log_cols + 1)) for _ in range(nq)]
    want_basis, want_sum = orc.induce(log_cols, rows, v, queries, alpha)
    qt = torch.tensor(queries, dtype=torch.int64, device="cuda")
    basis, s = hp.induce_sumcheck_poly(log_cols, to_dev(rows.reshape(-1)), v, qt, alpha)
    torch.cuda.synchronize()
    assert same(from_dev(basis), want_basis)
    assert f128_to_ints(from_dev(s))[0] == want_sum


def _armed(fn, out, nbytes):
    """Run fn (which enqueues a reduction writing `out`) with an armed read-back
    of nbytes; returns the host bytes the arm delivered."""
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    host = (ctypes.c_uint8 * nbytes)()
    nat.call("flk_read_arm", ctypes.cast(host, ctypes.c_void_p), nbytes)
    fn()
    nat.call("flk_read_armed", ctypes.c_void_p(out.data_ptr()), nbytes, hp._stream())
    return np.frombuffer(bytes(host), dtype=np.int64).reshape(-1, 2)


@pytest.mark.parametrize("logn", [2, 9, 16, 22])
def test_armed_read_back_equals_plain_read(orc, dev, logn):
    """The round messages read back through flk_read_arm/flk_read_armed (the
    reduction's last work-group stores them into host memory) equal the device
    outputs and the oracle, for one work-group and for thousands; the same
    calls, repeated, leave the stream's completion counter usable."""
    _log(f"armed read-back logn={logn}")
    rng = random.Random(7 * logn)
    f, b = orc.fill(500 + logn, 1 << logn), orc.fill(600 + logn, 1 << logn)
    df, db = to_dev(f), to_dev(b)
    for rep in range(3):
        out2 = torch.full((2, 2), -1, dtype=torch.int64, device="cuda")
        h = _armed(lambda: hp.round_msg(df, db, out=out2), out2, 32)
        want = orc.round_msg(f, b)
        assert same(from_dev(torch.from_numpy(h.copy()).cuda()), want), rep
        assert same(from_dev(out2), want), rep
        out3 = torch.full((3, 2), -1, dtype=torch.int64, device="cuda")
        h3 = _armed(lambda: hp.round_msg_and_eval(df, db, out=out3), out3, 48)
        assert same(from_dev(torch.from_numpy(h3.copy()).cuda()), orc.round_msg_and_eval(f, b)), rep
        if logn >= 2:
            r = rng.getrandbits(128)
            fo, bo, msg = orc.fold_and_msg(f, b, r)
            gm = torch.full((2, 2), -1, dtype=torch.int64, device="cuda")
            nf = torch.empty(((1 << logn) // 2, 2), dtype=torch.int64, device="cuda")
            nb = torch.empty_like(nf)
            hm = _armed(lambda: hp.fold_and_msg(df, db, r, f_out=nf, b_out=nb, out=gm), gm, 32)
            assert same(from_dev(torch.from_numpy(hm.copy()).cuda()), msg), rep
            assert same(from_dev(nf), fo) and same(from_dev(nb), bo), rep


def test_armed_read_back_falls_back_when_not_taken(orc, dev):
    """An arm whose size no reduction matches is not taken: flk_read_armed reads
    the device source the ordinary way.  Without an arm it is refused."""
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    f, b = orc.fill(701, 1 << 10), orc.fill(702, 1 << 10)
    out2 = torch.full((2, 2), -1, dtype=torch.int64, device="cuda")
    hp.round_msg(to_dev(f), to_dev(b), out=out2)
    torch.cuda.synchronize()
    # armed for 3 results, the round message has 2: read through the fallback
    host = (ctypes.c_uint8 * 48)()
    src = torch.cat([out2, torch.zeros((1, 2), dtype=torch.int64, device="cuda")])
    nat.call("flk_read_arm", ctypes.cast(host, ctypes.c_void_p), 48)
    hp.round_msg(to_dev(f), to_dev(b), out=out2)
    torch.cuda.synchronize()
    nat.call("flk_read_armed", ctypes.c_void_p(src.data_ptr()), 48, hp._stream())
    got = np.frombuffer(bytes(host), dtype=np.int64).reshape(-1, 2)
    assert (got == src.cpu().numpy()).all()
    rc = nat.load().flk_read_armed(ctypes.c_void_p(src.data_ptr()), 48, hp._stream())
    assert rc == nat.FLK_E_SHAPE
