"""Pins the oracle's Ligerito opening glue (oracle/open.c): the multi-proof
sibling indices (CF/merkle.rs:1067-1107) verified the way the reference's
verifier does (:1114-1191) against trees built by the oracle's Merkle code,
the single-position case equal to the ordinary authentication path
(merkle_proof, :960), and corruption rejected."""
import random

import numpy as np
import pytest


def _tree(orc, n, seed):
    data = np.random.default_rng(seed).integers(0, 256, n * 64, dtype=np.uint8).tobytes()
    return orc.merkle_tree(data, n)


@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024])
def test_multi_proof_verifies(orc, n):
    tree = _tree(orc, n, n)
    root = bytes(tree[-1])
    r = random.Random(n)
    for k in (1, 2, 5, 40, n):
        pos = [r.randrange(n) for _ in range(min(k, 3 * n))]
        pos += pos[: len(pos) // 3]  # repeats are allowed
        idx = orc.multi_proof_indices(n, pos)
        su = sorted(set(pos))
        proof = tree[idx] if idx else np.zeros((0, 32), np.uint8)
        assert orc.multi_verify(root, n, su, tree[su], proof)
        if len(idx):
            bad = proof.copy()
            bad[r.randrange(len(idx)), 0] ^= 1
            assert not orc.multi_verify(root, n, su, tree[su], bad)
            assert not orc.multi_verify(root, n, su, tree[su], proof[:-1])
        assert len(idx) <= len(su) * max(n.bit_length() - 1, 0)
        if len(su) == n:
            assert idx == []


def test_single_position_is_the_path(orc):
    n = 256
    tree = _tree(orc, n, 3)
    for i in (0, 1, 77, 255):
        idx = orc.multi_proof_indices(n, [i])
        path = orc.merkle_proof(tree, n, i)
        assert np.array_equal(tree[idx], path)


def test_siblings_share_parents(orc):
    # two siblings need no hash at the leaf level
    n = 16
    assert orc.multi_proof_indices(n, [4, 5])[0] == 16 + 3  # first emitted: level-1 sibling of parent 2
    assert len(orc.multi_proof_indices(n, [4, 5])) == 3
