"""GPU parity: lincheck partial fold and product-sumcheck rounds, ring-switch
folds and the zerocheck fold-at-z / multilinear rounds on gfx950 against the
oracle, bit-exact, through the C ABI.  The largest cases use the m = 26
(2^12-compression) shapes of the reference configuration."""
import random
import sys

import numpy as np
import pytest
import torch

import pyfield as pf
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import from_dev, same, to_dev
from oracle_ffi import f128_to_ints, ints_to_f128

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rand_bytes(seed, n):
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8)


@pytest.mark.parametrize("m,k_log", [(8, 3), (12, 4), (17, 14), (20, 14), (26, 14)])
def test_lc_partial_fold(orc, dev, m, k_log):
    _log(f"lc_partial_fold m={m} k_log={k_log}")
    rng = random.Random(m)
    zb = _rand_bytes(m, (1 << m) // 8)
    eq = orc.build_eq([rng.getrandbits(128) for _ in range(m - k_log)])
    want = orc.lc_partial_fold(zb, m, k_log, eq)
    got = hp.partial_fold_packed_z(torch.from_numpy(zb).cuda(), m, k_log, to_dev(eq))
    torch.cuda.synchronize()
    assert same(from_dev(got), want)


@pytest.mark.parametrize("logn", [1, 3, 9, 14, 18])
def test_lc_round_and_bind(orc, dev, logn):
    rng = random.Random(logn)
    comb, z = orc.fill(10 + logn, 1 << logn), orc.fill(20 + logn, 1 << logn)
    got = hp.lincheck_round_msg(to_dev(comb), to_dev(z))
    r = rng.getrandbits(128)
    t = to_dev(comb)
    bound = hp.sumcheck_bind_top_in_place(t, r)
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.lc_round_msg_msb(comb, z))
    assert same(from_dev(bound), orc.lc_bind_top(comb, r))


def test_lincheck_product_sumcheck_8_rounds(orc, dev):
    """The 8-round top-variable product sumcheck over 2^14 tables, round by round."""
    rng = random.Random(5)
    comb, z = orc.fill(1, 1 << 14), orc.fill(2, 1 << 14)
    tc, tz = to_dev(comb), to_dev(z)
    for rnd in range(8):
        msg = hp.lincheck_round_msg(tc, tz)
        want = orc.lc_round_msg_msb(comb, z)
        assert same(from_dev(msg), want), rnd
        r = rng.getrandbits(128)
        tc = hp.sumcheck_bind_top_in_place(tc, r).contiguous()
        tz = hp.sumcheck_bind_top_in_place(tz, r).contiguous()
        comb, z = orc.lc_bind_top(comb, r), orc.lc_bind_top(z, r)
    torch.cuda.synchronize()
    assert same(from_dev(tc), comb) and same(from_dev(tz), z)


@pytest.mark.parametrize("n", [1, 5, 256, 1000, 1 << 16, 1 << 19])
def test_rs_fold_1b_rows(orc, dev, n):
    _log(f"fold_1b_rows n={n}")
    z, eq = orc.fill(30 + n, n), orc.fill(40 + n, n)
    got = hp.fold_1b_rows(to_dev(z), to_dev(eq))
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.rs_fold_1b_rows(z, eq))


@pytest.mark.parametrize("n", [1, 7, 4096, 1 << 19])
def test_rs_fold_b128_elems(orc, dev, n):
    _log(f"fold_b128_elems n={n}")
    sfx, eqr = orc.fill(50 + n, n), orc.fill(60 + n, 128)
    got = hp.fold_b128_elems(to_dev(sfx), to_dev(eqr))
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.rs_fold_b128_elems(sfx, eqr))


def test_lagrange_weights_host(orc, dev):
    rng = random.Random(3)
    for k in (3, 6):
        z = rng.getrandbits(128)
        assert hp.lagrange_weights(k, z) == f128_to_ints(orc.zc_lagrange_weights(k, z))


@pytest.mark.parametrize("m,k_skip", [(6, 6), (9, 3), (12, 6), (20, 6), (26, 6), (14, 5)])
def test_zc_fold_at_z(orc, dev, m, k_skip):
    _log(f"fold_at_z m={m} k_skip={k_skip}")
    rng = random.Random(m * 10 + k_skip)
    bits = _rand_bytes(m + 100, (1 << m) // 8)
    w = orc.zc_lagrange_weights(k_skip, rng.getrandbits(128))
    got = hp.fold_at_z(torch.from_numpy(bits).cuda(), m, k_skip, to_dev(w))
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.zc_fold_at_z(bits, m, k_skip, w))


@pytest.mark.parametrize("logn", [1, 4, 12, 20])
def test_zc_round_pair(orc, dev, logn):
    _log(f"zc_round_pair logn={logn}")
    rng = random.Random(logn)
    a, b = orc.fill(70 + logn, 1 << logn), orc.fill(80 + logn, 1 << logn)
    eq_rest = orc.build_eq([rng.getrandbits(128) for _ in range(logn - 1)])
    got = hp.zerocheck_round_pair(to_dev(a), to_dev(b), to_dev(eq_rest))
    torch.cuda.synchronize()
    assert same(from_dev(got), orc.zc_round_pair(a, b, eq_rest))


def _fold_round(ta, tb, rho, eq_lo, eq_hi, hi_log, with_next=False):
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    n = ta.shape[0]
    out = hp.f128_empty(2)
    ao = bo = None
    r = None
    if rho is not None:
        r = nat.F128.from_int(rho)
        ao, bo = hp.f128_empty(n // 2), hp.f128_empty(n // 2)
    nxt = hp.f128_empty(max(1, (1 << hi_log) // 2)) if with_next else None
    p = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    nat.call("flk_zc_fold_round", p(ta), p(tb), n, ctypes.byref(r) if r is not None else None, p(eq_lo), p(eq_hi),
             hi_log, p(ao), p(bo), p(nxt), p(out), hp._stream())
    return out, ao, bo, nxt


@pytest.mark.parametrize("hi_log,fold", [(0, False), (0, True), (3, True), (5, False), (12, True)])
def test_zc_fold_round_matches_oracle(orc, dev, hi_log, fold):
    """The fused bind + round with a split eq equals partial_eval_lsb then round_pair
    against the full eq table (eq[x] = eq_lo[x mod 256] eq_hi[x >> 8]); the next
    round's eq_hi is eq_hi with its lowest coordinate summed out."""
    _log(f"zc_fold_round hi_log={hi_log} fold={fold}")
    rng = random.Random(hi_log * 2 + fold)
    n_pairs = 256 << hi_log
    n = 4 * n_pairs if fold else 2 * n_pairs
    a, b = orc.fill(90 + hi_log, n), orc.fill(91 + hi_log, n)
    pt = [rng.getrandbits(128) for _ in range(8 + hi_log)]
    rho = rng.getrandbits(128) if fold else None
    out, ao, bo, nxt = _fold_round(to_dev(a), to_dev(b), rho, hp.build_eq_table(pt[:8]), hp.build_eq_table(pt[8:]),
                                   hi_log, with_next=hi_log > 0)
    torch.cuda.synchronize()
    if fold:
        a, b = orc.partial_eval_lsb(a, [rho]), orc.partial_eval_lsb(b, [rho])
        assert same(from_dev(ao), a) and same(from_dev(bo), b)
    assert same(from_dev(out), orc.zc_round_pair(a, b, orc.build_eq(pt)))
    if hi_log > 0:
        assert same(from_dev(nxt), orc.build_eq(pt[9:]))


@pytest.mark.parametrize("m", [15, 16, 22])
def test_zc_fold_at_z_round_matches_oracle(orc, dev, m):
    """The first round straight from the bits: the tables equal the oracle's fold at
    z and the message its round pair against the full eq table."""
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    _log(f"zc_fold_at_z_round m={m}")
    rng = random.Random(m)
    abits, bbits = _rand_bytes(m + 7, (1 << m) // 8), _rand_bytes(m + 8, (1 << m) // 8)
    w = orc.zc_lagrange_weights(6, rng.getrandbits(128))
    hi_log = m - 15
    pt = [rng.getrandbits(128) for _ in range(8 + hi_log)]
    ta, tb = torch.from_numpy(abits).cuda(), torch.from_numpy(bbits).cuda()
    ao, bo, out = hp.f128_empty(1 << (m - 6)), hp.f128_empty(1 << (m - 6)), hp.f128_empty(2)
    nxt = hp.f128_empty(max(1, (1 << hi_log) // 2))
    lo, hi, wd = hp.build_eq_table(pt[:8]), hp.build_eq_table(pt[8:]), to_dev(w)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    nat.call("flk_zc_fold_at_z_round", p(ta), p(tb), m, p(wd), p(lo), p(hi), p(ao), p(bo),
             p(nxt) if hi_log else None, p(out), hp._stream())
    torch.cuda.synchronize()
    a, b = orc.zc_fold_at_z(abits, m, 6, w), orc.zc_fold_at_z(bbits, m, 6, w)
    assert same(from_dev(ao), a) and same(from_dev(bo), b)
    assert same(from_dev(out), orc.zc_round_pair(a, b, orc.build_eq(pt)))
    if hi_log:
        assert same(from_dev(nxt), orc.build_eq(pt[9:]))


@pytest.mark.parametrize("m", [15, 16, 22])
def test_zc_fold_at_z_round_message_only(orc, dev, m):
    """Round 0 without the tables (a_out = b_out = NULL): the eq_hi factor rides in
    the b side's scaled weights, and the message still equals the oracle's round
    pair over the fold at z against the full eq table."""
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    _log(f"zc_fold_at_z_round message-only m={m}")
    rng = random.Random(100 + m)
    abits, bbits = _rand_bytes(m + 17, (1 << m) // 8), _rand_bytes(m + 18, (1 << m) // 8)
    w = orc.zc_lagrange_weights(6, rng.getrandbits(128))
    hi_log = m - 15
    pt = [rng.getrandbits(128) for _ in range(8 + hi_log)]
    ta, tb = torch.from_numpy(abits).cuda(), torch.from_numpy(bbits).cuda()
    out = hp.f128_empty(2)
    nxt = hp.f128_empty(max(1, (1 << hi_log) // 2))
    lo, hi, wd = hp.build_eq_table(pt[:8]), hp.build_eq_table(pt[8:]), to_dev(w)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    nat.call("flk_zc_fold_at_z_round", p(ta), p(tb), m, p(wd), p(lo), p(hi), None, None,
             p(nxt) if hi_log else None, p(out), hp._stream())
    torch.cuda.synchronize()
    a, b = orc.zc_fold_at_z(abits, m, 6, w), orc.zc_fold_at_z(bbits, m, 6, w)
    assert same(from_dev(out), orc.zc_round_pair(a, b, orc.build_eq(pt)))
    if hi_log:
        assert same(from_dev(nxt), orc.build_eq(pt[9:]))


@pytest.mark.parametrize("m", [16, 17, 23])
def test_zc_fold_at_z_bind_round_matches_oracle(orc, dev, m):
    """Round 1 from the bits: the written tables equal the oracle's fold at z bound
    by rho (partial_eval_lsb), and the message its round pair over them."""
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    _log(f"zc_fold_at_z_bind_round m={m}")
    rng = random.Random(200 + m)
    abits, bbits = _rand_bytes(m + 27, (1 << m) // 8), _rand_bytes(m + 28, (1 << m) // 8)
    w = orc.zc_lagrange_weights(6, rng.getrandbits(128))
    rho = rng.getrandbits(128)
    hi_log = m - 16
    pt = [rng.getrandbits(128) for _ in range(8 + hi_log)]
    ta, tb = torch.from_numpy(abits).cuda(), torch.from_numpy(bbits).cuda()
    ao, bo, out = hp.f128_empty(1 << (m - 7)), hp.f128_empty(1 << (m - 7)), hp.f128_empty(2)
    nxt = hp.f128_empty(max(1, (1 << hi_log) // 2))
    lo, hi, wd = hp.build_eq_table(pt[:8]), hp.build_eq_table(pt[8:]), to_dev(w)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    r = nat.F128.from_int(rho)
    nat.call("flk_zc_fold_at_z_bind_round", p(ta), p(tb), m, p(wd), ctypes.byref(r), p(lo), p(hi), p(ao), p(bo),
             p(nxt) if hi_log else None, p(out), hp._stream())
    torch.cuda.synchronize()
    a = orc.partial_eval_lsb(orc.zc_fold_at_z(abits, m, 6, w), [rho])
    b = orc.partial_eval_lsb(orc.zc_fold_at_z(bbits, m, 6, w), [rho])
    assert same(from_dev(ao), a) and same(from_dev(bo), b)
    assert same(from_dev(out), orc.zc_round_pair(a, b, orc.build_eq(pt)))
    if hi_log:
        assert same(from_dev(nxt), orc.build_eq(pt[9:]))


def test_zc_fold_round_rejects_shapes(dev):
    import ctypes
    from flock_challenge_multi_amd import _native as nat
    t = hp.f128_empty(1024)
    p = ctypes.c_void_p(t.data_ptr())
    lib = nat.load()
    assert lib.flk_zc_fold_round(p, p, 1000, None, p, p, 0, None, None, None, p, None) == nat.FLK_E_SHAPE
    assert lib.flk_zc_fold_round(p, p, 1024, ctypes.byref(nat.F128.from_int(3)), p, p, 0, None, None, None, p,
                                 None) == nat.FLK_E_SHAPE
    assert lib.flk_zc_fold_at_z_round(p, p, 14, p, p, p, p, p, None, p, None) == nat.FLK_E_SHAPE
    assert lib.flk_zc_fold_at_z_round(p, p, 16, p, p, p, p, None, None, p, None) == nat.FLK_E_SHAPE
    r = nat.F128.from_int(5)
    assert lib.flk_zc_fold_at_z_bind_round(p, p, 15, p, ctypes.byref(r), p, p, p, p, None, p, None) == nat.FLK_E_SHAPE
    assert lib.flk_zc_fold_at_z_bind_round(p, p, 16, p, None, p, p, p, p, None, p, None) == nat.FLK_E_SHAPE
    assert lib.flk_is_disabled() == 0


def test_zerocheck_multilinear_rounds_chain(orc, dev):
    """fold at z, then round message + LSB bind per round with the eq table of the
    remaining challenges, as the prover chains them (CF/zerocheck.rs:1748-1781)."""
    rng = random.Random(17)
    m, k_skip = 16, 6
    nv = m - k_skip
    abits, bbits = _rand_bytes(1, (1 << m) // 8), _rand_bytes(2, (1 << m) // 8)
    z = rng.getrandbits(128)
    w = orc.zc_lagrange_weights(k_skip, z)
    a = orc.zc_fold_at_z(abits, m, k_skip, w)
    b = orc.zc_fold_at_z(bbits, m, k_skip, w)
    ta = hp.fold_at_z(torch.from_numpy(abits).cuda(), m, k_skip, to_dev(w))
    tb = hp.fold_at_z(torch.from_numpy(bbits).cuda(), m, k_skip, to_dev(w))
    r = [rng.getrandbits(128) for _ in range(nv)]
    for i in range(nv):
        eq_rest_pt = r[i + 1:]
        eq_rest = orc.build_eq(eq_rest_pt)
        msg = hp.zerocheck_round_pair(ta, tb, hp.build_eq_table(eq_rest_pt))
        assert same(from_dev(msg), orc.zc_round_pair(a, b, eq_rest)), i
        rho = rng.getrandbits(128)
        ta, tb = hp.partial_eval_lsb(ta, [rho]), hp.partial_eval_lsb(tb, [rho])
        a, b = orc.partial_eval_lsb(a, [rho]), orc.partial_eval_lsb(b, [rho])
    torch.cuda.synchronize()
    assert same(from_dev(ta), a) and same(from_dev(tb), b)


def _fixed_inner_challenges():
    import json
    import os
    fix = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "phi8_table.json")))["table"]
    phi = [int(a, 16) | (int(b, 16) << 64) for a, b in fix]
    small = [phi[0xF7], phi[0x53], phi[0xB5]]
    medium = [pf.mul(g, pf.inv(1 ^ g)) for g in (2, 4, 16, 256)]
    return small + medium


@pytest.mark.parametrize("m", [13, 14, 16, 20, 22])
def test_zc_round1_matches_oracle(orc, dev, m):
    """Round 1 for the protocol challenge layout: r[0..6) unused by round 1,
    r[6..13) fixed, r[13..m) random; against round1_naive on the oracle."""
    _log(f"zc_round1 m={m}")
    rng = random.Random(m)
    nb = (1 << m) // 8
    a, b, c = (_rand_bytes(1000 * m + i, nb) for i in range(3))
    r_outer = [rng.getrandbits(128) for _ in range(m - 13)]
    r = [rng.getrandbits(128) for _ in range(6)] + _fixed_inner_challenges() + r_outer
    want_ab, want_c = orc.zc_round1(a, b, c, m, 6, r)
    ta, tb, tc = (torch.from_numpy(x).cuda() for x in (a, b, c))
    got_ab, got_c = hp.zerocheck_round1(ta, tb, tc, m, r_outer)
    torch.cuda.synchronize()
    assert same(from_dev(got_ab), want_ab)
    assert same(from_dev(got_c), want_c)


def test_zc_round1_honest_witness_vanishes_on_s(orc, dev):
    """For a witness with c = a AND b the zerocheck polynomial P_AB + P_C has a
    zero on S; its Lambda evaluations from the device equal the oracle's, and the
    C part alone is F2-linear in c."""
    m = 14
    nb = (1 << m) // 8
    a, b = _rand_bytes(7, nb), _rand_bytes(8, nb)
    c = a & b
    rng = random.Random(3)
    r_outer = [rng.getrandbits(128)]
    r = [0] * 6 + _fixed_inner_challenges() + r_outer
    want_ab, want_c = orc.zc_round1(a, b, c, m, 6, r)
    got_ab, got_c = hp.zerocheck_round1(*(torch.from_numpy(x).cuda() for x in (a, b, c)), m, r_outer)
    torch.cuda.synchronize()
    assert same(from_dev(got_ab), want_ab) and same(from_dev(got_c), want_c)
