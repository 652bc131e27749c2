"""Pins the oracle's lincheck, ring-switch and zerocheck reductions.

Closed forms (independent of the oracle's loops):
  * lincheck partial fold = MLE of z over the outer variables at x_outer,
    evaluated per inner index from unpacked bits (CF/lincheck.rs:2220-2232 doc);
  * product-sumcheck round: q(0) + q(1) = claim and binding the top variable
    at r gives sum of products = q(r) (verifier identity, CF/lincheck.rs:5985-5998);
  * fold_1b_rows: sum_r x^r s_hat_v[r] = sum_i z_i eq_i (the ring-switching
    repacking identity, bits in the polynomial basis);
  * fold_b128_elems with eq_r[b] = c x^b is multiplication by c;
  * Lagrange weights: L_i(phi8(j)) = delta_ij, sum_i L_i(z) = 1, and they
    interpolate a random degree < 2^k_skip polynomial;
  * zerocheck round pair: (1 + r0) G(0) + r0 G(1) = sum_x eq(r, x) a_x b_x and
    G(rho) = G(0)(1+rho) + G(1) rho + G(inf) rho (1+rho) after binding at rho
    (CF/zerocheck.rs:1748-1781).
"""
import random

import numpy as np
import pytest

import pyfield as pf
from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128


def _bits_of(zb, idx):
    return (zb[idx >> 3] >> (idx & 7)) & 1


def test_lc_partial_fold_is_outer_mle(orc):
    rng = random.Random(1)
    m, k_log = 8, 3
    k, n_outer = 1 << k_log, 1 << (m - k_log)
    zb = np.frombuffer(rng.randbytes((1 << m) // 8), dtype=np.uint8).copy()
    x_outer = [rng.getrandbits(128) for _ in range(m - k_log)]
    eq = orc.build_eq(x_outer)
    got = f128_to_ints(orc.lc_partial_fold(zb, m, k_log, eq))
    eqi = f128_to_ints(eq)
    for i in range(k):
        want = 0
        for o in range(n_outer):
            s, r = o // 8, o % 8
            if (zb[s * k + i] >> r) & 1:
                want ^= pf.eq_entry(x_outer, o)
        assert got[i] == want
    assert eqi == [pf.eq_entry(x_outer, o) for o in range(n_outer)]


def test_lc_round_and_bind_identity(orc):
    rng = random.Random(2)
    n = 32
    comb, z = orc.fill(3, n), orc.fill(4, n)
    c, zz = f128_to_ints(comb), f128_to_ints(z)
    e1, einf = f128_to_ints(orc.lc_round_msg_msb(comb, z))
    h = n // 2
    q0 = 0
    for i in range(h):
        q0 ^= pf.mul(c[i], zz[i])
    assert e1 == _dot(c[h:], zz[h:])
    claim = _dot(c, zz)
    assert q0 ^ e1 == claim
    r = rng.getrandbits(128)
    bc, bz = orc.lc_bind_top(comb, r), orc.lc_bind_top(z, r)
    assert f128_to_ints(bc) == [c[i] ^ pf.mul(r, c[i + h] ^ c[i]) for i in range(h)]
    q_r = pf.mul(q0, 1 ^ r) ^ pf.mul(e1, r) ^ pf.mul(einf, pf.mul(r, 1 ^ r))
    assert q_r == _dot(f128_to_ints(bc), f128_to_ints(bz))


def _dot(a, b):
    s = 0
    for x, y in zip(a, b):
        s ^= pf.mul(x, y)
    return s


def test_fold_1b_rows_repacking_identity(orc):
    z, eq = orc.fill(5, 100), orc.fill(6, 100)
    s = f128_to_ints(orc.rs_fold_1b_rows(z, eq))
    lhs = 0
    for r in range(128):
        lhs ^= pf.mul(1 << r, s[r])
    assert lhs == _dot(f128_to_ints(z), f128_to_ints(eq))
    # and the direct per-bit definition
    zi, ei = f128_to_ints(z), f128_to_ints(eq)
    for r in (0, 1, 63, 64, 127):
        want = 0
        for a, e in zip(zi, ei):
            if (a >> r) & 1:
                want ^= e
        assert s[r] == want


def test_fold_b128_is_linear_map(orc):
    rng = random.Random(7)
    c = rng.getrandbits(128)
    eq_r = ints_to_f128([pf.mul(c, 1 << b) for b in range(128)])
    sfx = orc.fill(8, 50)
    got = f128_to_ints(orc.rs_fold_b128_elems(sfx, eq_r))
    assert got == [pf.mul(c, v) for v in f128_to_ints(sfx)]


def test_lagrange_weights(orc):
    phi = [int(e["lo"]) | (int(e["hi"]) << 64) for e in orc.phi8_table()]
    k = 3
    for j in range(1 << k):
        w = f128_to_ints(orc.zc_lagrange_weights(k, phi[j]))
        assert w == [1 if i == j else 0 for i in range(1 << k)]
    rng = random.Random(9)
    z = rng.getrandbits(128)
    w = f128_to_ints(orc.zc_lagrange_weights(k, z))
    acc = 0
    for x in w:
        acc ^= x
    assert acc == 1
    coeffs = [rng.getrandbits(128) for _ in range(1 << k)]

    def poly(x):
        r, p = 0, 1
        for c in coeffs:
            r ^= pf.mul(c, p)
            p = pf.mul(p, x)
        return r
    interp = 0
    for i in range(1 << k):
        interp ^= pf.mul(w[i], poly(phi[i]))
    assert interp == poly(z)


def test_fold_at_z_definition(orc):
    rng = random.Random(10)
    m, k_skip = 9, 6
    bits = np.frombuffer(rng.randbytes((1 << m) // 8), dtype=np.uint8).copy()
    w = orc.zc_lagrange_weights(k_skip, rng.getrandbits(128))
    got = f128_to_ints(orc.zc_fold_at_z(bits, m, k_skip, w))
    wi = f128_to_ints(w)
    for x in range(1 << (m - k_skip)):
        want = 0
        for s in range(1 << k_skip):
            if _bits_of(bits, x * (1 << k_skip) + s):
                want ^= wi[s]
        assert got[x] == want


def test_zc_round_pair_verifier_identity(orc):
    rng = random.Random(11)
    nv = 4
    n = 1 << nv
    a, b = orc.fill(12, n), orc.fill(13, n)
    r = [rng.getrandbits(128) for _ in range(nv)]
    eq_rest = orc.build_eq(r[1:])
    g1, ginf = f128_to_ints(orc.zc_round_pair(a, b, eq_rest))
    ai, bi = f128_to_ints(a), f128_to_ints(b)
    claim = 0
    for x in range(n):
        claim ^= pf.mul(pf.eq_entry(r, x), pf.mul(ai[x], bi[x]))
    eqr = f128_to_ints(eq_rest)
    g0 = 0
    for x in range(n // 2):
        g0 ^= pf.mul(eqr[x], pf.mul(ai[2 * x], bi[2 * x]))
    assert pf.mul(1 ^ r[0], g0) ^ pf.mul(r[0], g1) == claim
    rho = rng.getrandbits(128)
    ab = f128_to_ints(orc.partial_eval_lsb(a, [rho]))
    bb = f128_to_ints(orc.partial_eval_lsb(b, [rho]))
    g_rho = pf.mul(g0, 1 ^ rho) ^ pf.mul(g1, rho) ^ pf.mul(ginf, pf.mul(rho, 1 ^ rho))
    want = 0
    for x in range(n // 2):
        want ^= pf.mul(eqr[x], pf.mul(ab[x], bb[x]))
    assert g_rho == want


def _interp_eval(xs, ys, z):
    """Lagrange interpolation over GF(2^8) through (xs, ys), evaluated at z."""
    acc = 0
    for i, (xi, yi) in enumerate(zip(xs, ys)):
        num, den = 1, 1
        for j, xj in enumerate(xs):
            if j != i:
                num = pf.mul8(num, z ^ xj)
                den = pf.mul8(den, xi ^ xj)
        # den^-1 = den^254
        inv = 1
        for _ in range(254):
            inv = pf.mul8(inv, den)
        acc ^= pf.mul8(yi, pf.mul8(num, inv))
    return acc


@pytest.mark.parametrize("k", [1, 2, 3, 4])
def test_gf8_lde_is_polynomial_extension(orc, k):
    """inverse NTT on S = {0..2^k-1} then forward NTT on Lambda = {2^k..} evaluates the
    unique degree < 2^k interpolant of the S values at the Lambda points
    (CF/zerocheck/univariate_skip.rs:1-20)."""
    rng = random.Random(k)
    ell = 1 << k
    vals = [rng.randrange(256) for _ in range(ell)]
    coeff = orc.gf8_inverse(k, 0, vals)
    assert orc.gf8_forward(k, 0, coeff) == vals  # round trip on S
    lam = orc.gf8_forward(k, ell, coeff)
    xs = list(range(ell))
    assert lam == [_interp_eval(xs, vals, ell + i) for i in range(ell)]


def test_zc_round1_closed_form(orc):
    """round1_naive against the defining sum with the LDE by interpolation."""
    rng = random.Random(21)
    m, k = 6, 3
    ell = 1 << k
    nb = (1 << m) // 8
    a, b, c = (np.frombuffer(rng.randbytes(nb), np.uint8).copy() for _ in range(3))
    r = [rng.getrandbits(128) for _ in range(m)]
    pab, pc = orc.zc_round1(a, b, c, m, k, r)
    phi = [int(e["lo"]) | (int(e["hi"]) << 64) for e in orc.phi8_table()]
    want_ab, want_c = [0] * ell, [0] * ell
    xs = list(range(ell))
    for x in range(1 << (m - k)):
        bits = lambda arr: [int((arr[(x * ell + s) >> 3] >> ((x * ell + s) & 7)) & 1) for s in range(ell)]
        av, bv, cv = bits(a), bits(b), bits(c)
        e = pf.eq_entry(r[k:], x)
        for i in range(ell):
            al, bl, cl = (_interp_eval(xs, v, ell + i) for v in (av, bv, cv))
            want_ab[i] ^= pf.mul(e, phi[pf.mul8(al, bl)])
            want_c[i] ^= pf.mul(e, phi[cl])
    assert f128_to_ints(pab) == want_ab
    assert f128_to_ints(pc) == want_c


@pytest.mark.parametrize("m,threads", [(6, 1), (9, 3), (16, 8)])
def test_zc_round1_mt_equals_naive(orc, m, threads):
    """The CPU baseline's threaded, bucketed round 1 equals round1_naive bit for bit,
    on random bits and on all-ones / all-zeros columns."""
    rng = random.Random(m * 7 + threads)
    nb = (1 << m) // 8
    a, b, c = (np.frombuffer(rng.randbytes(nb), np.uint8).copy() for _ in range(3))
    a[: nb // 4] = 0xFF
    b[nb // 8: nb // 2] = 0
    r = [rng.getrandbits(128) for _ in range(m)]
    want = orc.zc_round1(a, b, c, m, 6, r)
    got = orc.zc_round1_mt(a, b, c, m, 6, r, threads)
    assert all(f128_to_ints(g) == f128_to_ints(w) for g, w in zip(got, want))
