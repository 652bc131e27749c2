"""Pins the oracle's ring-switch proving and claim combination
(oracle/ring_switch.c) the way the reference's own ring-switch tests do
(CF/pcs/ring_switch.rs, CF = crates/flock-core/src/):

* the oracle's transcript byte log equals the library's host FsChallenger
  (itself pinned against a byte-stream model over the published BLAKE3
  vectors) on random scripts;
* transpose_is_involution (:5222-5250);
* claim_check_recovers_zhat_skip (:5106-5135): the verifier's claim check on
  the prover's s_hat_v equals the zerocheck claim z_hat(z_skip, x_outer)
  computed directly from the bits of a Boolean witness;
* dp24_identity_holds (:5176-5196): <packed witness, rs_eq_ind> equals the
  BaseFold claim, here for the gamma-combined basis and target of 1-3 claims;
* verify_rejects_mutated_proof (:5199-5220).
Host code only: runs without a GPU."""
import random

import numpy as np
import pytest

from flock_challenge_multi_amd import hotpath as hp
from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128


def rand_f128(r):
    return r.getrandbits(128)


def packed_from_bits(bits):
    """pack_witness: bit j -> element j >> 7, polynomial bit j & 127."""
    by = np.packbits(bits.astype(np.uint8), bitorder="little")
    return by.view(F128_DT).copy()


def inner(orc, a, b):
    p = f128_to_ints(orc.mul_many(a, b))
    acc = 0
    for v in p:
        acc ^= v
    return acc


@pytest.mark.parametrize("seed", range(4))
def test_transcript_log_matches_library(orc, seed):
    r = random.Random(seed)
    dom = bytes(r.randrange(256) for _ in range(r.randrange(30)))
    a, b = orc.fs(dom), hp.FsChallenger(dom)
    for _ in range(40):
        op = r.randrange(5)
        if op == 0:
            l = bytes(r.randrange(256) for _ in range(r.randrange(40)))
            a.observe_label(l), b.observe_label(l)
        elif op == 1:
            v = rand_f128(r)
            a.observe_f128(v), b.observe_f128(v)
        elif op == 2:
            vs = [rand_f128(r) for _ in range(r.choice([0, 1, 7, 128]))]
            a.observe_f128_slice(vs), b.observe_f128_slice(vs)
        elif op == 3:
            assert a.sample_f128() == b.sample_f128()
        else:
            n = r.choice([0, 1, 7, 20])
            assert a.sample_f128_vec(n) == b.sample_f128_vec(n)
    assert a.state_digest() == b.state_digest()


def test_transpose_is_involution(orc):
    r = random.Random(5)
    for _ in range(8):
        v = [rand_f128(r) for _ in range(128)]
        once = orc.rs_transpose(v)
        assert orc.rs_transpose(once) == v
    unit = [1 << i for i in range(128)]
    assert orc.rs_transpose(unit) == unit  # the identity matrix is symmetric
    assert orc.rs_transpose([0] * 128) == [0] * 128
    assert orc.rs_transpose([2**128 - 1] * 128) == [2**128 - 1] * 128
    # one set bit: bit b of element i moves to bit i of element b
    v = [0] * 128
    v[3] = 1 << 100
    w = [0] * 128
    w[100] = 1 << 3
    assert orc.rs_transpose(v) == w


def zhat_skip(orc, bits, m, z_skip, x_outer):
    """z_hat(z_skip, x_outer) = sum_{i_outer} eq(x_outer, i_outer) sum_{i_skip} lambda_{i_skip}(z_skip)
    z[64 i_outer + i_skip], straight from the bits."""
    lam = f128_to_ints(orc.zc_lagrange_weights(6, z_skip))
    eq_outer = orc.build_eq(x_outer)
    inner_vals = []
    for io in range(1 << (m - 6)):
        acc = 0
        for s in range(64):
            if bits[64 * io + s]:
                acc ^= lam[s]
        inner_vals.append(acc)
    return inner(orc, eq_outer, ints_to_f128(inner_vals))


@pytest.mark.parametrize("m", [8, 9, 10])
def test_claim_check_recovers_zhat_skip(orc, m):
    r = random.Random(0xAA7 + m)
    rng = np.random.default_rng(m)
    bits = rng.integers(0, 2, 1 << m, dtype=np.uint8)
    z_skip = rand_f128(r)
    x_outer = [rand_f128(r) for _ in range(m - 6)]
    want = zhat_skip(orc, bits, m, z_skip, x_outer)
    shv, _, _, _ = orc.rs_prove_batched(orc.fs(b"flock-test-v0"), packed_from_bits(bits), [x_outer], basis=False)
    assert orc.rs_claim_check(z_skip, x_outer[0], shv[0]) == want
    # verify_rejects_mutated_proof: one flipped bit of s_hat_v breaks the check
    bad = list(shv[0])
    bad[0] ^= 1
    assert orc.rs_claim_check(z_skip, x_outer[0], bad) != want


@pytest.mark.parametrize("m,claims", [(8, 1), (9, 2), (10, 3), (11, 2)])
def test_dp24_identity_combined(orc, m, claims):
    r = random.Random(m * 10 + claims)
    rng = np.random.default_rng(m)
    z = packed_from_bits(rng.integers(0, 2, 1 << m, dtype=np.uint8))
    xs = [[rand_f128(r) for _ in range(m - 6)] for _ in range(claims)]
    shv, cl, ga, b, target = orc.pcs_combined_basis_and_target(orc.fs(b"flock-test-v0"), z, xs)
    assert inner(orc, z, b) == target
    acc = 0
    for g, c in zip(ga, cl):
        acc ^= orc.mul(g, c)
    assert acc == target
    # one claim at a time: <z, rs_eq_ind_k> = claim_k, with rs_eq_ind_k = fold_b128(eq(x_k[1..]), eq(r''_k))
    t = orc.fs(b"flock-test-v0")
    t.observe_label(b"flock-pcs-open-batch-v0")
    for k in range(claims):
        t.observe_label(b"flock-ring-switch-v0")
        t.observe_f128_slice(shv[k])
        eq_r = orc.build_eq(t.sample_f128_vec(7))
        rs = orc.rs_fold_b128_elems(orc.build_eq(xs[k][1:]), eq_r)
        assert inner(orc, z, rs) == cl[k]
        assert shv[k] == f128_to_ints(orc.rs_fold_1b_rows(z, orc.build_eq(xs[k][1:])))
    assert [t.sample_f128() for _ in range(claims)] == ga
