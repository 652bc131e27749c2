"""GPU parity: the Ligerito OOD loop through the C ABI (flk_ligerito_ood) against
the oracle (oracle/ligerito_ood.c): OOD values, messages, the glued basis, the
running target and the transcript state, bit for bit; at 2^25 elements the
sumcheck invariant <f, b'> = T' on the device."""
import random

import numpy as np
import pytest
import torch

from flock_challenge_multi_amd import _native as nat
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import from_dev, same, to_dev
from oracle_ffi import f128_to_ints

pytestmark = pytest.mark.gpu


def inner(orc, a, b):
    acc = 0
    for v in f128_to_ints(orc.mul_many(a, b)):
        acc ^= v
    return acc


@pytest.mark.parametrize("log_n,samples", [(1, 1), (2, 2), (5, 1), (10, 2), (14, 3), (19, 2)])
def test_ood_matches_oracle(orc, dev, log_n, samples):
    f = orc.fill(log_n + 1, 1 << log_n)
    b = orc.fill(log_n + 50, 1 << log_n)
    t0 = random.Random(log_n).getrandbits(128)
    fd, bd = to_dev(f), to_dev(b)
    sp = hp.SumcheckProver(fd, bd, t0)
    t_dev, t_orc = hp.FsChallenger(b"flock-test-v0"), orc.fs(b"flock-test-v0")
    for t in (t_dev, t_orc):
        t.observe_label(b"L1:root")
    ys = sp.ood(t_dev, samples)
    torch.cuda.synchronize()
    w_b, w_t, w_ys, w_msgs = orc.ligerito_ood(t_orc, f, b, samples, t0)
    assert ys == w_ys and sp.t_r == w_t and sp.transcript == w_msgs
    assert same(from_dev(sp.combined_basis), w_b)
    assert t_dev.state_digest() == t_orc.state_digest()


def test_ood_zero_samples_and_shapes(orc, dev):
    f = to_dev(orc.fill(1, 8))
    b = to_dev(orc.fill(2, 8))
    sp = hp.SumcheckProver(f, b, 5)
    t = hp.FsChallenger(b"x")
    d0 = t.state_digest()
    assert sp.ood(t, 0) == [] and sp.t_r == 5 and t.state_digest() == d0
    import ctypes
    lib = nat.load()
    T = nat.F128()
    ys, ms = (nat.F128 * 1)(), (nat.F128 * 2)()
    assert lib.flk_ligerito_ood(t._h, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(b.data_ptr()), 0, 1,
                                ctypes.byref(T), ys, ms, None) == 1
    assert lib.flk_ligerito_ood(t._h, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(b.data_ptr()), 3, 65,
                                ctypes.byref(T), ys, ms, None) == 1
    assert t.state_digest() == d0


def test_ood_invariant_full_size(orc, dev):
    """2^25 elements (the m = 32 L0 basis length): <f, b> = T before gives <f, b'> = T' after two samples."""
    n = 1 << 25
    g = torch.Generator(device="cuda").manual_seed(5)
    f = torch.randint(-2**62, 2**62, (n, 2), dtype=torch.int64, device="cuda", generator=g)
    b = torch.randint(-2**62, 2**62, (n, 2), dtype=torch.int64, device="cuda", generator=g)
    t0 = f128_to_ints(from_dev(hp.round_msg_and_eval(f, b)))[2]
    sp = hp.SumcheckProver(f, b, t0)
    ys = sp.ood(hp.FsChallenger(b"flock-test-v0"), 2)
    t1 = f128_to_ints(from_dev(hp.round_msg_and_eval(f, sp.combined_basis)))[2]
    assert len(ys) == 2 and t1 == sp.t_r
