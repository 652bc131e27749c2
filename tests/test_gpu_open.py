"""GPU parity: Ligerito opening glue through the C ABI — the opened rows
gathered from a device codeword, and the Merkle multi-proof of a device tree
against the oracle's sibling indices, verified against the root like the
reference's verifier (CF/merkle.rs:1114-1191)."""
import random

import numpy as np
import pytest
import torch

from flock_challenge_multi_amd import _native as nat
from flock_challenge_multi_amd import hotpath as hp
from gpu_util import from_dev, same, to_dev

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n_rows,row_len", [(1, 1), (64, 8), (1 << 17, 64)])
def test_gather_rows(orc, dev, n_rows, row_len):
    cw = orc.fill(n_rows + row_len, n_rows * row_len)
    r = random.Random(n_rows)
    pos = [r.randrange(n_rows) for _ in range(min(218, 3 * n_rows))]
    got = hp.gather_rows(to_dev(cw), row_len, pos)
    torch.cuda.synchronize()
    want = np.concatenate([cw[p * row_len:(p + 1) * row_len] for p in pos])
    assert same(from_dev(got), want)


@pytest.mark.parametrize("n", [2, 16, 1 << 10, 1 << 17])
def test_multi_proof_matches_oracle(orc, dev, n):
    data = torch.randint(0, 256, (n * 64,), dtype=torch.uint8, device="cuda")
    tree = hp.merkle_tree(data, n)
    torch.cuda.synchronize()
    host_tree = tree.cpu().numpy()
    root = bytes(host_tree[-1])
    r = random.Random(n)
    for k in (1, 3, 106, 218):
        pos = [r.randrange(n) for _ in range(k)]
        got = hp.merkle_multi_proof(tree, n, pos).cpu().numpy()
        idx = orc.multi_proof_indices(n, pos)
        assert np.array_equal(got, host_tree[idx] if idx else np.zeros((0, 32), np.uint8))
        su = sorted(set(pos))
        assert orc.multi_verify(root, n, su, host_tree[su], got)


def test_open_shape_errors(dev):
    import ctypes
    t = hp.hashes_empty(31)
    pos = (ctypes.c_uint64 * 2)(3, 99)
    k = ctypes.c_size_t()
    assert dev.flk_merkle_multi_proof(t.data_ptr(), 16, pos, 2, t.data_ptr(), 64, ctypes.byref(k), None) == nat.FLK_E_SHAPE
    assert dev.flk_merkle_multi_proof(t.data_ptr(), 12, pos, 1, t.data_ptr(), 64, ctypes.byref(k), None) == nat.FLK_E_SHAPE
    pos1 = (ctypes.c_uint64 * 1)(3)
    assert dev.flk_merkle_multi_proof(t.data_ptr(), 16, pos1, 1, t.data_ptr(), 2, ctypes.byref(k), None) == nat.FLK_E_SHAPE
    cw = hp.f128_empty(64)
    assert dev.flk_gather_rows(cw.data_ptr(), 8, 8, pos, 2, cw.data_ptr(), None) == nat.FLK_E_SHAPE
    assert dev.flk_is_disabled() == 0
