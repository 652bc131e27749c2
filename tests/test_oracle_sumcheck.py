"""Pins the oracle's eq tables, folds and Ligerito sumcheck reductions.

Closed forms used (independent of the oracle's loops):
  * eq table entry i = prod_j (bit_j(i) ? r_j : 1 + r_j), per entry
    (CF/lincheck.rs:2075-2080 doc);
  * folding all variables LSB-first = MLE evaluation sum_i e_i eq(r, i);
  * round message (u_0, u_2) satisfies the verifier's RoundQuad identity
    (CF/pcs/ligerito.rs:4929-4956): for g(X) = sum_j f_j(X) b_j(X),
    g(0) = u_0, leading coefficient = u_2, and g(r) = sum f'b' after a fold;
  * induce_sumcheck_poly basis = sum_q alpha_q X_j(x_q) with the LCH basis
    X_j(x) = prod_{k in bits(j)} What_k(x) evaluated by brute force, and the
    enforced sum = sum_q alpha_q <row_q, eq(v)>.
"""
import random

import numpy as np
import pytest

import pyfield as pf
from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128


def test_eq_table_per_entry(orc):
    rng = random.Random(11)
    for d in [0, 1, 3, 6]:
        pt = [rng.getrandbits(128) for _ in range(d)]
        eq = f128_to_ints(orc.build_eq(pt))
        assert eq == [pf.eq_entry(pt, i) for i in range(1 << d)]


def test_full_fold_is_mle_evaluation(orc):
    rng = random.Random(12)
    k = 5
    ev = orc.fill(3, 1 << k)
    rs = [rng.getrandbits(128) for _ in range(k)]
    got = f128_to_ints(orc.partial_eval_lsb(ev, rs))[0]
    vals = f128_to_ints(ev)
    want = 0
    for i, v in enumerate(vals):
        want ^= pf.mul(v, pf.eq_entry(rs, i))
    assert got == want
    # partial folds compose
    mid = orc.partial_eval_lsb(ev, rs[:2])
    assert f128_to_ints(orc.partial_eval_lsb(mid, rs[2:]))[0] == want


def _line_sum(f, b, x):
    """sum_j f_j(x) b_j(x) with f_j(X) = f[2j] + X (f[2j] + f[2j+1])."""
    s = 0
    for j in range(len(f) // 2):
        fx = f[2 * j] ^ pf.mul(x, f[2 * j] ^ f[2 * j + 1])
        bx = b[2 * j] ^ pf.mul(x, b[2 * j] ^ b[2 * j + 1])
        s ^= pf.mul(fx, bx)
    return s


def test_round_msg_satisfies_roundquad_identity(orc):
    rng = random.Random(13)
    n = 32
    f_a, b_a = orc.fill(21, n), orc.fill(22, n)
    f, b = f128_to_ints(f_a), f128_to_ints(b_a)
    u0, u2 = f128_to_ints(orc.round_msg(f_a, b_a))
    T = 0
    for x, y in zip(f, b):
        T ^= pf.mul(x, y)
    assert u0 == _line_sum(f, b, 0)
    assert T ^ u0 == _line_sum(f, b, 1)  # g(1) = T + g(0) in characteristic 2
    for _ in range(4):
        r = rng.getrandbits(128)
        quad = u0 ^ pf.mul(r, T ^ u2) ^ pf.mul(pf.mul(r, r), u2)  # c + r b + r^2 a
        assert quad == _line_sum(f, b, r)


def test_round_msg_and_eval(orc):
    f_a, b_a = orc.fill(31, 64), orc.fill(32, 64)
    u0, u2, y = f128_to_ints(orc.round_msg_and_eval(f_a, b_a))
    assert [u0, u2] == f128_to_ints(orc.round_msg(f_a, b_a))
    T = 0
    for x, z in zip(f128_to_ints(f_a), f128_to_ints(b_a)):
        T ^= pf.mul(x, z)
    assert y == T


def test_fold_and_msg_is_fold_then_msg(orc):
    rng = random.Random(14)
    n = 64
    f_a, b_a = orc.fill(41, n), orc.fill(42, n)
    r = rng.getrandbits(128)
    fo, bo, msg = orc.fold_and_msg(f_a, b_a, r)
    f, b = f128_to_ints(f_a), f128_to_ints(b_a)
    assert f128_to_ints(fo) == [f[2 * j] ^ pf.mul(r, f[2 * j] ^ f[2 * j + 1]) for j in range(n // 2)]
    assert f128_to_ints(bo) == [b[2 * j] ^ pf.mul(r, b[2 * j] ^ b[2 * j + 1]) for j in range(n // 2)]
    assert f128_to_ints(msg) == f128_to_ints(orc.round_msg(fo, bo))
    # folding is the partial evaluation at r
    assert (fo == orc.partial_eval_lsb(f_a, [r])).all()


def test_glue(orc):
    b, bn = orc.fill(51, 40), orc.fill(52, 40)
    alpha = 0x1234_5678_9ABC_DEF0_0FED_CBA9_8765_4321
    got = f128_to_ints(orc.glue(b, bn, alpha))
    assert got == [x ^ pf.mul(alpha, y) for x, y in zip(f128_to_ints(b), f128_to_ints(bn))]


def test_eval_sk_at_vks_matches_subspace_polys(orc):
    # sks_vks[i] = W_i(x^i) (unnormalised subspace poly at the next basis vector)
    log_n = 6
    sk = f128_to_ints(orc.eval_sk_at_vks(log_n))
    for i in range(log_n + 1):
        w = 1
        for v in range(1 << i):
            w = pf.mul(w, (1 << i) ^ v)
        assert sk[i] == w, i


def test_induce_sumcheck_poly_closed_form(orc):
    rng = random.Random(15)
    log_cols, n_v, n_alpha = 5, 3, 3
    nq = 6
    rows = orc.fill(61, nq * (1 << n_v)).reshape(nq, 1 << n_v)
    v = [rng.getrandbits(128) for _ in range(n_v)]
    alpha = [rng.getrandbits(128) for _ in range(n_alpha)]
    queries = [rng.randrange(1 << (log_cols + 2)) for _ in range(nq)]
    basis, s = orc.induce(log_cols, rows, v, queries, alpha)
    aq = [pf.eq_entry(alpha, q) for q in range(nq)]
    eqv = [pf.eq_entry(v, k) for k in range(1 << n_v)]
    want_sum = 0
    for qi in range(nq):
        dot = 0
        for k, x in enumerate(f128_to_ints(rows[qi])):
            dot ^= pf.mul(x, eqv[k])
        want_sum ^= pf.mul(dot, aq[qi])
    assert s == want_sum
    want = [0] * (1 << log_cols)
    for qi, q in enumerate(queries):
        what = [orc.subspace_poly_hat(k, q) for k in range(log_cols)]
        for j in range(1 << log_cols):
            x = aq[qi]
            for k in range(log_cols):
                if (j >> k) & 1:
                    x = pf.mul(x, what[k])
            want[j] ^= x
    assert f128_to_ints(basis) == want
