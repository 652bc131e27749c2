"""Pins the oracle's additive NTT (CF/ntt/additive_ntt_f128.rs).

Closed form: the forward transform maps novel-basis coefficients c_j to the
evaluations P(z) = sum_j c_j prod_{k in bits(j)} What_k(z) at z = position
index (as a field element), where What_k is the normalised subspace
polynomial — computed here by brute force over the span, independently of the
twiddle recurrence.  Also: the reference's literal twiddle KATs
(:8104-8109 twiddle(0,0)=0, :9913-9924 deepest two layers have hi == 0),
forward/inverse round trip and linearity (:8038-8101), and the RS-encode
definition (zero-padded full transform, CF/pcs/commit.rs:1838-1877).
"""
import numpy as np
import pytest

import pyfield as pf
from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128


@pytest.mark.parametrize("dim", [1, 2, 3, 5, 8, 12])
def test_twiddle_kats(orc, dim):
    tw = f128_to_ints(orc.twiddles(dim))
    assert len(tw) == (1 << dim) - 1
    for layer in range(dim):
        assert tw[(1 << layer) - 1] == 0  # twiddle(l, 0) = 0
    for layer in range(max(dim - 2, 0), dim):  # deepest two layers fit in lo
        for b in range(1 << layer):
            assert tw[(1 << layer) - 1 + b] >> 64 == 0
    # deepest layer is the span of x^1..x^(dim-1): twiddle(dim-1, b) = 2b
    for b in range(1 << (dim - 1)):
        assert tw[(1 << (dim - 1)) - 1 + b] == 2 * b


def test_twiddles_are_a_span(orc):
    # each layer is an F2-linear span: t(b1 ^ b2) = t(b1) + t(b2)
    dim = 9
    tw = f128_to_ints(orc.twiddles(dim))
    for layer in range(dim):
        base = (1 << layer) - 1
        for b1 in range(0, 1 << layer, 3):
            for b2 in range(0, 1 << layer, 5):
                assert tw[base + (b1 ^ b2)] == tw[base + b1] ^ tw[base + b2]


def _novel_eval(orc, coeffs, z, log_d):
    what = [orc.subspace_poly_hat(k, z) for k in range(log_d)]
    acc = 0
    for j, c in enumerate(coeffs):
        x = c
        for k in range(log_d):
            if (j >> k) & 1:
                x = pf.mul(x, what[k])
        acc ^= x
    return acc


@pytest.mark.parametrize("log_d,num_ntts", [(1, 1), (3, 1), (4, 2), (5, 1), (4, 4)])
def test_forward_is_novel_basis_evaluation(orc, log_d, num_ntts):
    data = orc.fill(0x51 + log_d, (1 << log_d) * num_ntts)
    out = f128_to_ints(orc.forward_interleaved(log_d, data, num_ntts))
    vals = f128_to_ints(data)
    for lane in range(num_ntts):
        coeffs = vals[lane::num_ntts]
        for pos in range(1 << log_d):
            assert out[pos * num_ntts + lane] == _novel_eval(orc, coeffs, pos, log_d), (lane, pos)


def test_subspace_poly_hat_definition(orc):
    # What_i(x^i) = 1, What_i vanishes on span(x^0..x^{i-1})
    for i in range(6):
        assert orc.subspace_poly_hat(i, 1 << i) == 1
        for v in range(1 << i):
            assert orc.subspace_poly_hat(i, v) == 0


@pytest.mark.parametrize("log_d", [1, 4, 9])
def test_forward_inverse_roundtrip(orc, log_d):
    data = orc.fill(99 + log_d, 1 << log_d)
    fwd = orc.forward_interleaved(log_d, data, 1)
    back = orc.inverse(log_d, fwd)
    assert (back == data).all()


def test_linearity(orc):
    a = orc.fill(1, 256 * 8)
    b = orc.fill(2, 256 * 8)
    ab = a.copy()
    ab["lo"] ^= b["lo"]
    ab["hi"] ^= b["hi"]
    fa, fb, fab = (orc.forward_interleaved(8, x, 8) for x in (a, b, ab))
    assert (fa["lo"] ^ fb["lo"] == fab["lo"]).all() and (fa["hi"] ^ fb["hi"] == fab["hi"]).all()


@pytest.mark.parametrize("log_d,r,num_ntts", [(6, 1, 4), (8, 2, 2), (7, 3, 8), (5, 5, 2)])
def test_rs_encode_is_zero_padded_transform_and_from_layer(orc, log_d, r, num_ntts):
    n_msg = (1 << (log_d - r)) * num_ntts
    msg = orc.fill(7 * log_d + r, n_msg)
    cw = orc.rs_encode(log_d, msg, n_msg << r, num_ntts)
    padded = np.zeros(n_msg << r, dtype=F128_DT)
    padded[:n_msg] = msg
    assert (cw == orc.forward_interleaved(log_d, padded, num_ntts)).all()
    # the first r layers are copies: replicate then start at layer r (:2785-2830)
    rep = np.tile(msg, 1 << r)
    assert (cw == orc.forward_interleaved(log_d, rep, num_ntts, start_layer=r)).all()
