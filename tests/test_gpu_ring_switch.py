"""GPU parity: ring-switch proving and the PCS claim combination through the C
ABI (flk_rs_prove_batched, flk_pcs_combined_basis_and_target) against the
oracle (oracle/ring_switch.c): s_hat_v, the BaseFold claims, gamma, the
combined basis, the target and the transcript state afterwards, bit for bit;
at the 2^32-bit witness (2^25 packed elements, two claims) the DP24 identity
<z, b> = target over the whole basis and sampled basis entries rebuilt on the
host."""
import random

import numpy as np
import pytest
import torch

from flock_challenge_multi_amd import _native as nat
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import from_dev, same, to_dev
from oracle_ffi import F128_DT, f128_to_ints, ints_to_f128

pytestmark = pytest.mark.gpu


def rand_f128(r):
    return r.getrandbits(128)


def prefix(t, r):
    """Some transcript history before the ring switch."""
    t.observe_label(b"flock-r1cs-v0")
    t.observe_f128(r)
    t.sample_f128()


@pytest.mark.parametrize("log_len,claims", [(0, 1), (1, 2), (3, 1), (5, 2), (5, 3), (6, 4), (6, 5),
                                            (9, 2), (12, 9), (14, 2), (20, 2), (20, 5)])
def test_combined_basis_matches_oracle(orc, dev, log_len, claims):
    r = random.Random(log_len * 100 + claims)
    z = orc.fill(log_len + 7 * claims, 1 << log_len)
    xs = [[rand_f128(r) for _ in range(log_len + 1)] for _ in range(claims)]
    seed = rand_f128(r)
    t_dev, t_orc = hp.FsChallenger(b"flock-test-v0"), orc.fs(b"flock-test-v0")
    prefix(t_dev, seed), prefix(t_orc, seed)
    shv, cl, ga, b, tgt = hp.compute_combined_basis_and_target(to_dev(z), xs, t_dev)
    torch.cuda.synchronize()
    w_shv, w_cl, w_ga, w_b, w_tgt = orc.pcs_combined_basis_and_target(t_orc, z, xs)
    assert shv == w_shv and cl == w_cl and ga == w_ga and tgt == w_tgt
    assert same(from_dev(b), w_b)
    assert t_dev.state_digest() == t_orc.state_digest()
    assert t_dev.sample_f128() == t_orc.sample_f128()


@pytest.mark.parametrize("want_basis", [True, False])
def test_prove_batched_matches_oracle(orc, dev, want_basis):
    r = random.Random(7)
    log_len, claims = 10, 3
    z = orc.fill(99, 1 << log_len)
    xs = [[rand_f128(r) for _ in range(log_len + 1)] for _ in range(claims)]
    t_dev, t_orc = hp.FsChallenger(b"d"), orc.fs(b"d")
    shv, cl, ga, b = hp.prove_batched(to_dev(z), xs, t_dev, want_basis=want_basis)
    w_shv, w_cl, w_ga, w_b = orc.rs_prove_batched(t_orc, z, xs)
    assert shv == w_shv and cl == w_cl and ga == w_ga
    if want_basis:
        assert same(from_dev(b), w_b)
    else:
        assert b is None
    assert t_dev.state_digest() == t_orc.state_digest()


def test_special_witnesses(orc, dev):
    """All-zero and all-one witnesses, and zero / one opening coordinates."""
    log_len = 6
    ones = np.zeros(1 << log_len, F128_DT)
    ones["lo"], ones["hi"] = np.uint64(2**64 - 1), np.uint64(2**64 - 1)
    for z in (np.zeros(1 << log_len, F128_DT), ones):
        for xs in ([[0] * (log_len + 1)], [[1] * (log_len + 1), [2**128 - 1] * (log_len + 1)]):
            t_dev, t_orc = hp.FsChallenger(b""), orc.fs(b"")
            shv, cl, ga, b, tgt = hp.compute_combined_basis_and_target(to_dev(z), xs, t_dev)
            w = orc.pcs_combined_basis_and_target(t_orc, z, xs)
            assert (shv, cl, ga, tgt) == (w[0], w[1], w[2], w[4]) and same(from_dev(b), w[3])


def test_shape_errors_leave_transcript_untouched(orc, dev):
    z = to_dev(orc.fill(1, 16))
    t = hp.FsChallenger(b"x")
    before = t.state_digest()
    with pytest.raises(ValueError):
        hp.compute_combined_basis_and_target(z, [[1] * 4], t)  # needs 5 coordinates
    with pytest.raises(ValueError):
        hp.compute_combined_basis_and_target(z, [], t)
    import ctypes
    F = nat.F128
    shv, cl, ga, tgt = (F * 128)(), (F * 1)(), (F * 1)(), F()
    xs = (F * 5)()
    lib = nat.load()
    assert lib.flk_pcs_combined_basis_and_target(t._h, ctypes.c_void_p(z.data_ptr()), 4, xs, 0, shv, cl, ga,
                                                 ctypes.c_void_p(z.data_ptr()), ctypes.byref(tgt), None) == 1
    assert lib.flk_pcs_combined_basis_and_target(t._h, ctypes.c_void_p(z.data_ptr()), 4, xs, 65, shv, cl, ga,
                                                 ctypes.c_void_p(z.data_ptr()), ctypes.byref(tgt), None) == 1
    assert lib.flk_pcs_combined_basis_and_target(t._h, ctypes.c_void_p(z.data_ptr()), 4, xs, 1, shv, cl, ga,
                                                 None, ctypes.byref(tgt), None) == 1
    assert t.state_digest() == before


def eq_entry(orc, point, i):
    """eq(point, i) = prod_j (bit_j(i) ? point_j : 1 + point_j)."""
    acc = 1
    for j, p in enumerate(point):
        acc = orc.mul(acc, p if (i >> j) & 1 else p ^ 1)
    return acc


def test_full_size_two_claims(orc, dev):
    """2^25 packed elements (a 2^32-bit witness), two claims: the DP24 identity
    <z, b> = target over the whole basis on the device, and basis entries
    rebuilt on the host from the replayed transcript."""
    log_len = 25
    n = 1 << log_len
    r = random.Random(32)
    g = torch.Generator(device="cuda").manual_seed(32)
    z = torch.randint(-2**62, 2**62, (n, 2), dtype=torch.int64, device="cuda", generator=g)
    xs = [[rand_f128(r) for _ in range(log_len + 1)] for _ in range(2)]
    t = hp.FsChallenger(b"flock-test-v0")
    shv, cl, ga, b, tgt = hp.compute_combined_basis_and_target(z, xs, t)
    msg = hp.round_msg_and_eval(z, b)
    torch.cuda.synchronize()
    assert f128_to_ints(from_dev(msg))[2] == tgt
    # replay the transcript on the oracle with the device's s_hat_v: the same gammas
    t_orc = orc.fs(b"flock-test-v0")
    t_orc.observe_label(b"flock-pcs-open-batch-v0")
    eq_rs = []
    for k in range(2):
        t_orc.observe_label(b"flock-ring-switch-v0")
        t_orc.observe_f128_slice(shv[k])
        eq_rs.append(orc.build_eq(t_orc.sample_f128_vec(7)))
    assert [t_orc.sample_f128() for _ in range(2)] == ga
    assert t_orc.state_digest() == t.state_digest()
    host_b = from_dev(b)
    for i in [0, 1, 2, 12345, n // 2 + 77, n - 1]:
        want = 0
        for k in range(2):
            eg = orc.mul_many(eq_rs[k], ints_to_f128([ga[k]] * 128))
            want ^= f128_to_ints(orc.rs_fold_b128_elems(ints_to_f128([eq_entry(orc, xs[k][1:], i)]), eg))[0]
        assert f128_to_ints(host_b[i:i + 1])[0] == want
