"""Pins the oracle's Ligerito OOD loop (oracle/ligerito_ood.c) by what the
mathematics fixes, with the transcript replayed independently:

* y is the multilinear extension of f at the sampled point z: the LSB-first
  partial evaluation of f over all of z (eval_mle_lsb, CF/pcs/ligerito.rs:1177-1182;
  the oracle's partial_eval_lsb is pinned separately);
* (u_0, u_2) is the round message of (f, eq(z)): with g(1) = y + u_0, the
  quadratic g(X) = u_0 + (u_0 + g(1) + u_2) X + u_2 X^2 equals
  sum fold(f, r) fold(eq(z), r) at a random r;
* the glue keeps the sumcheck invariant: if T = <f, b> before, T' = <f, b'>
  after, with b' = b + beta eq(z) for the replayed beta.
Host code only."""
import random

import numpy as np
import pytest

from oracle_ffi import f128_to_ints, ints_to_f128


def inner(orc, a, b):
    acc = 0
    for v in f128_to_ints(orc.mul_many(a, b)):
        acc ^= v
    return acc


@pytest.mark.parametrize("log_n,samples", [(2, 1), (3, 2), (6, 1), (9, 3)])
def test_ood_loop(orc, log_n, samples):
    r = random.Random(log_n * 10 + samples)
    f = orc.fill(log_n, 1 << log_n)
    b = orc.fill(log_n + 100, 1 << log_n)
    t0 = inner(orc, f, b)
    fs = orc.fs(b"flock-test-v0")
    fs.observe_label(b"root")
    b2, t2, ys, msgs = orc.ligerito_ood(fs, f, b, samples, t0)
    assert t2 == inner(orc, f, b2)
    # replay: the points, the values, the messages and the betas
    rp = orc.fs(b"flock-test-v0")
    rp.observe_label(b"root")
    bb = b.copy()
    for s in range(samples):
        z = rp.sample_f128_vec(log_n)
        y = f128_to_ints(orc.partial_eval_lsb(f, z))[0]
        assert ys[s] == y
        eq = orc.build_eq(z)
        u0, u2 = msgs[s]
        g1 = y ^ u0
        x = r.getrandbits(128)
        fo, eo, _ = orc.fold_and_msg(f, eq, x)
        want = u0 ^ orc.mul(u0 ^ g1 ^ u2, x) ^ orc.mul(u2, orc.mul(x, x))
        assert inner(orc, fo, eo) == want
        for v in (y, u0, u2):
            rp.observe_f128(v)
        beta = rp.sample_f128()
        bb = ints_to_f128([p ^ q for p, q in zip(f128_to_ints(bb), f128_to_ints(orc.mul_many(eq, ints_to_f128([beta] * len(eq)))))])
    assert f128_to_ints(b2) == f128_to_ints(bb)
    assert rp.state_digest() == fs.state_digest()
