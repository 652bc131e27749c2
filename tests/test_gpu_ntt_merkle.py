"""GPU parity: additive NTT / RS encode, BLAKE3 Merkle tree and PCS commit on
gfx950 against the oracle, bit-exact, through the C ABI.

Shapes cover the reference's own test geometries (CF/pcs/commit.rs:1838-1877:
(m, r, log_batch) in (10,1,1),(12,1,2),(12,2,1),(14,2,3)), the recursive
Ligerito levels (8 lanes, rates 1/4..1/64) and the full standalone commit of
2^22 coefficients (64 lanes, rate 1/2: 2^17 x 1 KiB leaves).
"""
import ctypes
import sys

import numpy as np
import pytest
import torch

import flock_challenge_multi_amd._native as nat
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import bytes_from_dev, from_dev, same, to_dev

pytestmark = pytest.mark.gpu


def _log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("log_d,r,num_ntts", [
    (1, 1, 1), (4, 1, 1), (5, 2, 2), (6, 1, 4), (7, 3, 8), (8, 2, 8), (9, 1, 64),
    (10, 4, 8), (11, 1, 16), (12, 1, 64), (12, 6, 8), (5, 5, 4), (13, 2, 32), (10, 1, 128),
])
def test_rs_encode_matches_oracle(orc, dev, log_d, r, num_ntts):
    _log(f"rs_encode log_d={log_d} r={r} lanes={num_ntts}")
    n_msg = (1 << (log_d - r)) * num_ntts
    msg = orc.fill(1000 * log_d + 10 * r + num_ntts, n_msg)
    want = orc.rs_encode(log_d, msg, n_msg << r, num_ntts)
    ntt = hp.AdditiveNttF128.standard(log_d)
    cw = hp.f128_empty(n_msg << r)
    cw.fill_(-1)  # stale contents must be overwritten
    ntt.rs_encode_interleaved(to_dev(msg), cw, num_ntts)
    torch.cuda.synchronize()
    assert same(from_dev(cw), want)


@pytest.mark.parametrize("log_d,start,num_ntts", [(6, 0, 1), (8, 3, 4), (10, 0, 64), (9, 5, 8), (7, 7, 2)])
def test_forward_interleaved_in_place(orc, dev, log_d, start, num_ntts):
    _log(f"forward log_d={log_d} start={start} lanes={num_ntts}")
    data = orc.fill(77 + log_d, (1 << log_d) * num_ntts)
    want = orc.forward_interleaved(log_d, data, num_ntts, start)
    t = to_dev(data)
    hp.AdditiveNttF128.standard(log_d).forward_transform_interleaved(t, num_ntts, start)
    torch.cuda.synchronize()
    assert same(from_dev(t), want)


@pytest.mark.parametrize("dim,log_d,num_ntts", [(1, 1, 1), (5, 5, 1), (9, 9, 1), (12, 9, 1), (12, 12, 1),
                                              (10, 10, 8), (8, 6, 64), (11, 11, 2)])
def test_inverse_matches_oracle(orc, dev, dim, log_d, num_ntts):
    """inverse_transform (CF/ntt/additive_ntt_f128.rs:6322) per lane, against the oracle."""
    _log(f"inverse dim={dim} log_d={log_d} lanes={num_ntts}")
    data = orc.fill(300 + 7 * dim + log_d, (1 << log_d) * num_ntts)
    lanes = data.reshape(1 << log_d, num_ntts)
    want = np.stack([orc.inverse(dim, np.ascontiguousarray(lanes[:, c])) for c in range(num_ntts)], axis=1)
    t = to_dev(data)
    hp.AdditiveNttF128.standard(dim).inverse_transform_interleaved(t, num_ntts)
    torch.cuda.synchronize()
    assert same(from_dev(t), want.reshape(-1))


@pytest.mark.parametrize("log_d,num_ntts", [(17, 64), (20, 8), (14, 1)])
def test_forward_inverse_round_trip(orc, dev, log_d, num_ntts):
    """forward then inverse is the identity, and inverse then forward too, at the
    standalone commit's full shape (2^17 positions x 64 lanes)."""
    _log(f"round trip log_d={log_d} lanes={num_ntts}")
    g = torch.Generator(device="cuda")
    g.manual_seed(log_d)
    x = torch.randint(-(2**63), 2**63 - 1, ((1 << log_d) * num_ntts, 2), dtype=torch.int64, device="cuda", generator=g)
    ntt = hp.AdditiveNttF128.standard(log_d)
    y = x.clone()
    ntt.forward_transform_interleaved(y, num_ntts)
    assert not torch.equal(y, x)
    ntt.inverse_transform_interleaved(y, num_ntts)
    assert torch.equal(y, x)
    ntt.inverse_transform_interleaved(y, num_ntts)
    ntt.forward_transform_interleaved(y, num_ntts)
    assert torch.equal(y, x)


def test_forward_with_larger_dim(orc, dev):
    # a transform smaller than the NTT's domain uses the dim-sized twiddle table
    data = orc.fill(5, 64 * 4)
    want = orc.forward_interleaved(9, data, 4)
    t = to_dev(data)
    hp.AdditiveNttF128.standard(9).forward_transform_interleaved(t, 4)
    torch.cuda.synchronize()
    assert same(from_dev(t), want)


@pytest.mark.parametrize("leaf_size,n_leaves", [(16, 1), (16, 64), (32, 8), (64, 256), (128, 1024),
                                               (256, 32), (512, 4096), (1024, 2048), (1024, 1)])
def test_merkle_tree_matches_oracle(orc, dev, leaf_size, n_leaves):
    _log(f"merkle leaf={leaf_size} n={n_leaves}")
    rng = np.random.default_rng(leaf_size * 7 + n_leaves)
    data = rng.integers(0, 256, leaf_size * n_leaves, dtype=np.uint8)
    want = orc.merkle_tree(data.tobytes(), n_leaves)
    d = torch.from_numpy(data).cuda()
    tree = hp.merkle_tree(d, n_leaves)
    torch.cuda.synchronize()
    assert (bytes_from_dev(tree) == want).all()


def test_hash_leaves_and_pairs(orc, dev):
    rng = np.random.default_rng(3)
    data = rng.integers(0, 256, 128 * 64, dtype=np.uint8)
    d = torch.from_numpy(data).cuda()
    leaves = hp.hashes_empty(64)
    hp.hash_leaves(d, 128, leaves)
    parents = hp.hashes_empty(32)
    hp.hash_pairs_level(leaves, parents)
    torch.cuda.synchronize()
    want = orc.merkle_tree(data.tobytes(), 64)
    assert (bytes_from_dev(leaves) == want[:64]).all()
    assert (bytes_from_dev(parents) == want[64:96]).all()


@pytest.mark.parametrize("m,r,log_batch", [(10, 1, 1), (12, 1, 2), (12, 2, 1), (14, 2, 3), (20, 1, 6), (16, 3, 3)])
def test_commit_matches_full_ntt_oracle(orc, dev, m, r, log_batch):
    _log(f"commit m={m} r={r} log_batch={log_batch}")
    params = hp.PcsParams(m=m, log_inv_rate=r, log_batch_size=log_batch)
    z = orc.fill(0xFEED + m, 1 << params.log_msg_len())
    cw, tree, root = orc.commit(z, params.num_ntts(), r)
    root_t, pd = hp.commit(to_dev(z), params)
    torch.cuda.synchronize()
    assert same(from_dev(pd.codeword), cw)
    assert (bytes_from_dev(pd.merkle_tree) == tree).all()
    assert bytes(root_t.cpu().numpy().tobytes()) == root


def test_commit_root_via_c_abi_host_copy(orc, dev):
    z = orc.fill(42, 1 << 12)
    _, _, root = orc.commit(z, 8, 2)
    zt = to_dev(z)
    cw = hp.f128_empty((1 << 12) << 2)
    tree = hp.hashes_empty(2 * (1 << 11) - 1)
    out = (ctypes.c_uint8 * 32)()
    nat.call("flk_commit", ctypes.c_void_p(zt.data_ptr()), 9, 3, 2, nat.FLK_HASH_BLAKE3,
             ctypes.c_void_p(cw.data_ptr()), ctypes.c_void_p(tree.data_ptr()), out, None)
    assert bytes(out) == root


def test_standalone_commit_full_size(orc, dev):
    """2^22 uniform coefficients, 2^16 positions x 64 lanes, rate 1/2:
    codeword (128 MiB) and root bit-exact to the oracle."""
    _log("commit full size 2^22")
    z = orc.fill(0xC0FFEE, 1 << 22)
    cw, tree, root = orc.commit(z, 64, 1)
    pd = hp.ligero_commit(to_dev(z), 16, 6, 1)
    torch.cuda.synchronize()
    assert same(from_dev(pd.codeword), cw)
    assert bytes(pd.root().cpu().numpy().tobytes()) == root
    assert (bytes_from_dev(pd.merkle_tree) == tree).all()


def test_shape_refusals_do_not_latch(dev):
    lib = nat.load()
    t = hp.f128_empty(64)
    with pytest.raises(ValueError):
        hp.AdditiveNttF128.standard(3).rs_encode_interleaved(t, hp.f128_empty(96), 4)
    st = lib.flk_ntt_rs_encode_interleaved(6, ctypes.c_void_p(t.data_ptr()), 64, ctypes.c_void_p(t.data_ptr()), 128,
                                           4, None)
    assert st == nat.FLK_E_SHAPE  # overlapping / wrong-size buffers
    st = lib.flk_merkle_tree(ctypes.c_void_p(t.data_ptr()), 1024, 4, nat.FLK_HASH_SHA256,
                             ctypes.c_void_p(t.data_ptr()), None)
    assert st == nat.FLK_E_SHAPE
    assert lib.flk_is_disabled() == 0


def test_commit_from_host_buffers_via_c_abi(orc, dev):
    """The host-buffer route a Rust caller takes: flk_alloc + flk_copy_to_device,
    flk_commit, flk_copy_to_host of the codeword, flk_synchronize, flk_free."""
    lib = nat.load()
    z = orc.fill(77, 1 << 13)
    cw_want, tree_want, root_want = orc.commit(z, 8, 2)
    n_cw, n_pos = (1 << 13) << 2, (1 << 12)  # 8 lanes x 2^10 cols at rate 1/4 -> 2^12 positions
    ptrs = [ctypes.c_void_p() for _ in range(3)]
    for p, nbytes in zip(ptrs, (16 << 13, 16 * n_cw, 32 * (2 * n_pos - 1))):
        nat.call("flk_alloc", nbytes, ctypes.byref(p))
    try:
        nat.call("flk_copy_to_device", ptrs[0], z.ctypes.data_as(ctypes.c_void_p), 16 << 13, None)
        root = (ctypes.c_uint8 * 32)()
        nat.call("flk_commit", ptrs[0], 10, 3, 2, nat.FLK_HASH_BLAKE3, ptrs[1], ptrs[2], root, None)
        cw = np.zeros_like(cw_want)
        nat.call("flk_copy_to_host", cw.ctypes.data_as(ctypes.c_void_p), ptrs[1], 16 * n_cw, None)
        nat.call("flk_synchronize", None)
        assert bytes(root) == root_want
        assert same(cw, cw_want)
    finally:
        for p in ptrs:
            nat.call("flk_free", p)


@pytest.mark.parametrize("log_d,r,num_ntts", [(12, 1, 64), (11, 2, 8), (9, 1, 64)])
def test_every_autotuner_schedule_gives_the_same_codeword(orc, dev, log_d, r, num_ntts):
    """Each candidate pass schedule of the autotuner (flk_ntt_set_schedule) groups
    the same layers differently: the codeword equals the oracle's for all of them,
    and the inverse transform under a custom schedule undoes the forward one."""
    from flock_challenge_multi_amd import autotune
    lanes_log = num_ntts.bit_length() - 1
    n_layers = log_d - r
    n_msg = (1 << (log_d - r)) * num_ntts
    msg = orc.fill(7000 + log_d, n_msg)
    want = orc.rs_encode(log_d, msg, n_msg << r, num_ntts)
    ntt = hp.AdditiveNttF128.standard(log_d)
    try:
        for sched in autotune.candidates(n_layers):
            _log(f"schedule {sched} log_d={log_d} r={r} lanes={num_ntts}")
            autotune.install(n_layers, sched, lanes_log)
            assert autotune.current(n_layers, lanes_log) == sched
            cw = hp.f128_empty(n_msg << r)
            cw.fill_(-1)
            ntt.rs_encode_interleaved(to_dev(msg), cw, num_ntts)
            torch.cuda.synchronize()
            assert same(from_dev(cw), want), sched
        sched = autotune.candidates(log_d)[0]
        autotune.install(log_d, sched, lanes_log)
        data = to_dev(want)
        ntt.inverse_transform_interleaved(data, num_ntts)
        ntt.forward_transform_interleaved(data, num_ntts)
        torch.cuda.synchronize()
        assert same(from_dev(data), want)
    finally:
        autotune.install(n_layers, (), lanes_log)
        autotune.install(log_d, (), lanes_log)
