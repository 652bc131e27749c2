"""GPU: the recursive Ligerito prover (flock_challenge_multi_amd.ligerito,
recursive_prover_with_basis) on device data, checked by the verifier
restated over the oracle (tests/ligerito_verifier.py, the reference's
recursive_verifier_with_basis): honest proofs are accepted, and a flipped yr
value, sumcheck message, opened row, recursive root, OOD value, grind nonce
or target is rejected.  The L0 root also equals the oracle's commit root."""
import copy

import numpy as np
import pytest
import torch

from flock_challenge_multi_amd import hotpath as hp
from flock_challenge_multi_amd import ligerito as lig
from gpu_util import from_dev, to_dev
from ligerito_verifier import verify
from oracle_ffi import f128_to_ints

pytestmark = pytest.mark.gpu

CONFIGS = {
    "two_levels": (12, dict(initial_k=3, recursive_ks=[2, 2], log_inv_rates=[1, 2, 2], queries=[8, 6, 5],
                            grinding_bits=[2, 2, 2], fold_grinding_bits=[3, 2, 1], ood_samples=[0, 1, 1])),
    "three_levels": (14, dict(initial_k=4, recursive_ks=[3, 2, 2], log_inv_rates=[1, 2, 3, 2],
                              queries=[12, 9, 7, 5], grinding_bits=[3, 1, 0, 2], fold_grinding_bits=[4, 0, 2, 1],
                              ood_samples=[0, 2, 1, 0])),
}


def inner(orc, a, b):
    acc = 0
    for v in f128_to_ints(orc.mul_many(a, b)):
        acc ^= v
    return acc


def host_proof(p, cfg):
    def rows(op, k):
        return from_dev(op.opened_rows).reshape(-1, 1 << k)

    ks = [cfg["initial_k"]] + cfg["recursive_ks"]
    return {
        "initial_root": p.initial_root,
        "initial_proof": (rows(p.initial_proof, cfg["initial_k"]), p.initial_proof.merkle_proof.cpu().numpy()),
        "recursive_roots": list(p.recursive_roots),
        "recursive_proofs": [(rows(op, ks[i + 1]), op.merkle_proof.cpu().numpy())
                             for i, op in enumerate(p.recursive_proofs)],
        "yr": list(p.final_proof.yr),
        "final_rows": rows(p.final_proof, ks[-1]),
        "final_merkle": p.final_proof.merkle_proof.cpu().numpy(),
        "sumcheck_transcript": list(p.sumcheck_transcript),
        "grinding_nonces": list(p.grinding_nonces),
        "ood_values": list(p.ood_values),
        "fold_grinding_nonces": list(p.fold_grinding_nonces),
    }


def prove(orc, name, seed):
    log_n, cfg = CONFIGS[name]
    f, b = orc.fill(seed, 1 << log_n), orc.fill(seed + 1, 1 << log_n)
    fd, bd = to_dev(f), to_dev(b)
    k0, r0 = cfg["initial_k"], cfg["log_inv_rates"][0]
    pd = hp.ligero_commit(fd, log_n - k0, k0, r0)
    target = inner(orc, f, b)
    p = lig.recursive_prover_with_basis(lig.ProverConfig(**cfg), fd, bd, target, pd.codeword, pd.merkle_tree,
                                        hp.FsChallenger(b"flock-test-v0"))
    torch.cuda.synchronize()
    _, _, root = orc.commit(f, 1 << k0, r0)
    return cfg, p, b, target, root


@pytest.mark.parametrize("name", list(CONFIGS))
def test_prover_accepted_by_verifier(orc, dev, name):
    cfg, p, b, target, root = prove(orc, name, 7)
    assert p.initial_root == root
    hpf = host_proof(p, cfg)
    assert len(hpf["grinding_nonces"]) == len(cfg["recursive_ks"]) + 1
    assert len(hpf["ood_values"]) == sum(cfg["ood_samples"])
    assert verify(orc, cfg, hpf, b, target, root, orc.fs(b"flock-test-v0"))


def test_tampered_proofs_rejected(orc, dev):
    cfg, p, b, target, root = prove(orc, "two_levels", 11)
    good = host_proof(p, cfg)
    assert verify(orc, cfg, good, b, target, root, orc.fs(b"flock-test-v0"))

    def bad(mut):
        h = copy.deepcopy(good)
        mut(h)
        return not verify(orc, cfg, h, b, target, root, orc.fs(b"flock-test-v0"))

    def flip_row(h):
        rows = h["recursive_proofs"][0][0].copy()
        rows[1, 0]["lo"] ^= np.uint64(1)
        h["recursive_proofs"][0] = (rows, h["recursive_proofs"][0][1])

    assert bad(lambda h: h["yr"].__setitem__(0, h["yr"][0] ^ 1))
    assert bad(lambda h: h["sumcheck_transcript"].__setitem__(3, (h["sumcheck_transcript"][3][0] ^ 2,
                                                                  h["sumcheck_transcript"][3][1])))
    assert bad(flip_row)
    assert bad(lambda h: h["recursive_roots"].__setitem__(1, bytes(32)))
    assert bad(lambda h: h["ood_values"].__setitem__(0, h["ood_values"][0] ^ 1))
    assert bad(lambda h: h["grinding_nonces"].__setitem__(0, h["grinding_nonces"][0] + 1))
    assert bad(lambda h: h["fold_grinding_nonces"].__setitem__(1, h["fold_grinding_nonces"][1] + 1))
    assert not verify(orc, cfg, good, b, target ^ 1, root, orc.fs(b"flock-test-v0"))
    # the basis is not in the transcript: only the final <yr, combined basis> = T check can see it
    b2 = b.copy()
    b2[5]["hi"] ^= np.uint64(1)
    assert not verify(orc, cfg, good, b2, target, root, orc.fs(b"flock-test-v0"))


def test_cpp_mirror_matches(orc, dev, tmp_path):
    """The C++ host mirror's recursive_prover_with_basis (include/flock_host.hpp,
    run by tests/cpp/host_mirror_test) writes the same proof as the Python one
    on the same instance, and the restated verifier accepts it."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "host_mirror_test")
    out = tmp_path / "proof.txt"
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    fields = {}
    for line in out.read_text().splitlines():
        key, *vals = line.split(" ")
        fields[key] = vals
    cfg, p, b, target, root = prove(orc, "two_levels", 7)
    f128s = lambda k: [int(v, 16) for v in fields[k]]
    hashes = lambda k: [bytes.fromhex(v) for v in fields[k]]
    assert hashes("initial_root") == [p.initial_root] and hashes("recursive_roots") == p.recursive_roots
    assert f128s("yr") == p.final_proof.yr and f128s("ood_values") == p.ood_values
    msgs = f128s("sumcheck_transcript")
    assert list(zip(msgs[::2], msgs[1::2])) == p.sumcheck_transcript
    assert [int(v) for v in fields["grinding_nonces"]] == p.grinding_nonces
    assert [int(v) for v in fields["fold_grinding_nonces"]] == p.fold_grinding_nonces
    for key, op in (("initial", p.initial_proof), ("recursive", p.recursive_proofs[0]), ("final", p.final_proof)):
        assert f128s(f"{key}_rows") == f128_to_ints(from_dev(op.opened_rows))
        assert hashes(f"{key}_merkle") == [bytes(h) for h in op.merkle_proof.cpu().numpy()]
