"""GPU tests of the full BLAKE3 R1CS prover (flk_blake3_prove) and its protocol
stages (flk_zerocheck_prove, flk_lincheck_prove, flk_lincheck_comb), checked by
the restated verifier (tests/flock_verifier.py) and the oracle.

  configs 0 and 2   generate_compressions(12, seed) (m = 26) and (18, seed) (m = 32):
                    the proof is accepted, including the commitment root recomputed
                    on the CPU from the regenerated witness, and stays under 500,000 B
  oracle prover     the proof equals the oracle CPU prover's bytes (a padded m = 22
                    batch, config 0 at m = 26, config 2 at m = 32: the bench's first seed)
  determinism       the same inputs give the same bytes (smallest-nonce grinds), on
                    another stream too; the golden fixture is reproduced exactly
  stages            round 1's messages equal the oracle's round1_naive at the
                    replayed challenges; the final a, b evaluations equal the MLE of
                    the oracle's fold at z; the lincheck comb equals the oracle's
                    sparse fold; lincheck's output claim w equals the witness MLE
"""
import os

import numpy as np
import pytest
import torch

import flock_verifier as fv
from oracle_ffi import COMPRESSION_DT, OracleFs, f128_to_ints, ints_to_f128
from flock_challenge_multi_amd import hotpath as hp
from flock_challenge_multi_amd import prover as pv
from flock_challenge_multi_amd.inputs import generate_compressions

pytestmark = pytest.mark.gpu


def _blocks(log2, seed, n=None):
    c = generate_compressions(log2, seed)
    if n is not None:
        c = c[:n]
    return np.ascontiguousarray(c), torch.from_numpy(np.ascontiguousarray(c).view(np.int32)).cuda()


_SETUPS = {}


def setup_for(n_blocks):
    if n_blocks not in _SETUPS:
        _SETUPS[n_blocks] = pv.Blake3Setup(n_blocks)
    return _SETUPS[n_blocks]


@pytest.mark.parametrize("log2,seed", [(8, 7), (12, 0x00C0FFEEBEEFD15C), (18, 0xF1A5C0)],
                         ids=["m22", "config0_m26", "config2_m32"])
def test_proof_verifies(orc, log2, seed):
    print(f"case prove log2={log2} seed={seed:#x}", flush=True)
    comps, blocks = _blocks(log2, seed)
    s = setup_for(1 << log2)
    proof, stages = s.prove(blocks)
    assert s.m == 14 + log2
    assert len(proof) < fv.MAX_PROOF_BYTES
    assert len(proof) <= s._cap  # flk_blake3_proof_bound
    fv.verify_proof(orc, proof, comps.view(COMPRESSION_DT).reshape(-1), log2)


@pytest.mark.parametrize("log2,seed,n", [(8, 0xA11CE, 200), (12, 0x00C0FFEEBEEFD15C, None), (18, 0x5EED0000, None)],
                         ids=["m22_padded200", "config0_m26", "config2_m32"])
def test_proof_equals_oracle_prover(orc, log2, seed, n):
    """The device proof equals, byte for byte, the proof of the oracle's CPU prover
    (tests/oracle_prover.py) on the same compressions."""
    import oracle_prover as op
    print(f"case oracle-prover log2={log2} seed={seed:#x} n={n}", flush=True)
    comps, blocks = _blocks(log2, seed, n)
    proof, _ = setup_for(len(comps)).prove(blocks)
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    want, _ = op.prove(orc, comps.view(COMPRESSION_DT).reshape(-1), log2, threads=threads)
    assert proof == want


def test_padded_batch_verifies(orc):
    comps, blocks = _blocks(8, 99, n=200)  # 200 compressions in 256 slots
    s = setup_for(200)
    assert s.n_blocks_log == 8
    proof, _ = s.prove(blocks)
    fv.verify_proof(orc, proof, comps.view(COMPRESSION_DT).reshape(-1), 8)


def test_proof_is_deterministic_and_stream_independent():
    _, blocks = _blocks(8, 31)
    s = setup_for(256)
    p1, _ = s.prove(blocks)
    st = torch.cuda.Stream()
    with torch.cuda.stream(st):
        p2, _ = s.prove(blocks, stream=st)
    st.synchronize()
    p3, _ = s.prove(blocks)
    assert p1 == p2 == p3
    _, other = _blocks(8, 32)
    assert s.prove(other)[0] != p1


def test_arena_reuse_and_concurrent_proves():
    """Proves of one setup draw their device memory from the setup's arenas (no
    pool fallback, peak within capacity); two host threads proving on two streams
    at once take two arenas and write the same bytes as a lone prove."""
    import threading
    _, blocks = _blocks(8, 41)
    s = pv.Blake3Setup(256)
    want, _ = s.prove(blocks)
    st = s.arena_stats()
    assert st["n_arenas"] == 1 and st["fallbacks"] == 0 and 0 < st["peak"] <= st["capacity"]
    assert s.prove(blocks)[0] == want and s.arena_stats()["n_arenas"] == 1  # the arena is reused
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    outs = [None, None]

    def run(i):
        for _ in range(2):
            outs[i] = s.prove(blocks, stream=streams[i])[0]

    th = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join(timeout=120)
    assert outs == [want, want]
    st = s.arena_stats()
    assert 1 <= st["n_arenas"] <= 2 and st["fallbacks"] == 0


def test_reproduces_golden_fixture():
    _, blocks = _blocks(8, 0x123C)
    proof, _ = setup_for(256).prove(blocks)
    assert proof == open(os.path.join(os.path.dirname(__file__), "golden", "proof_m22_seed123c.bin"), "rb").read()


def test_domain_changes_the_proof(orc):
    comps, blocks = _blocks(8, 5)
    s = setup_for(256)
    proof, _ = s.prove(blocks, domain=b"another-domain")
    b = fv.parse_bundle(proof)
    fv.verify_r1cs(orc, b, domain=b"another-domain")
    with pytest.raises(fv.Reject):
        fv.verify_r1cs(orc, b)


def _witness(log2, seed):
    _, blocks = _blocks(log2, seed)
    z, a, b, stripe = hp.generate_witness_with_ab_packed_and_lincheck(blocks, log2)
    torch.cuda.synchronize()
    return z, a, b, stripe


@pytest.mark.parametrize("log2", [8, 12])
def test_zerocheck_prove_matches_oracle(orc, log2):
    m = 14 + log2
    z, a, b, _ = _witness(log2, 11 + log2)
    ch = hp.FsChallenger(b"zc-test")
    proof, claim = pv.zerocheck_prove(ch, a, b, z, m)
    # the verifier replays the transcript and accepts
    och = OracleFs(orc, b"zc-test")
    vp = {"round1_ab": proof["round1_ab"], "round1_c": proof["round1_c"], "rounds": proof["multilinear_rounds"],
          "final_a": proof["final_a_eval"], "final_b": proof["final_b_eval"], "final_c": proof["final_c_eval"]}
    vc = fv.zerocheck_verify(orc, m, vp, och)
    assert vc["z"] == claim["z"] and vc["mlv"] == claim["mlv_challenges"] and vc["r_rest"] == claim["r_rest"]
    # round 1 equals round1_naive at the replayed challenges (r_skip is not used by round 1)
    abytes, bbytes, zbytes = (t.cpu().numpy().tobytes() for t in (a, b, z))
    r = [0] * 6 + claim["r_rest"]
    pab, pc = orc.zc_round1(np.frombuffer(abytes, np.uint8), np.frombuffer(bbytes, np.uint8),
                            np.frombuffer(zbytes, np.uint8), m, 6, r)
    assert f128_to_ints(pab) == proof["round1_ab"] and f128_to_ints(pc) == proof["round1_c"]
    # the final evaluations are the MLEs of the oracle's fold at z
    w = orc.zc_lagrange_weights(6, claim["z"])
    for bits, fin in ((abytes, "final_a_eval"), (bbytes, "final_b_eval")):
        mlv = orc.zc_fold_at_z(np.frombuffer(bits, np.uint8), m, 6, w)
        assert f128_to_ints(orc.partial_eval_lsb(mlv, claim["mlv_challenges"]))[0] == proof[fin]


def test_lincheck_comb_matches_oracle(orc):
    import random
    rng = random.Random(5)
    r1cs = pv.R1cs.blake3(8)
    z_skip, alpha, beta = (rng.getrandbits(128) for _ in range(3))
    xr = [rng.getrandbits(128) for _ in range(8)]
    got = r1cs.lincheck_comb(z_skip, xr, alpha, beta)
    torch.cuda.synchronize()
    want = fv.lincheck_comb(orc, alpha, fv.quirky_eq(orc, z_skip, xr))
    wi = f128_to_ints(want)
    wi[fv.CONST_POS] ^= beta
    gi = f128_to_ints(got.cpu().numpy().view(want.dtype).reshape(-1))
    assert gi == wi


def test_lincheck_prove_matches_oracle(orc):
    log2 = 8
    m = 14 + log2
    z, a, b, stripe = _witness(log2, 77)
    ch = hp.FsChallenger(b"lc-test")
    zp, zc = pv.zerocheck_prove(ch, a, b, z, m)
    x_ab = [zc["z"]] + zc["mlv_challenges"]
    r1cs = pv.R1cs.blake3(log2)
    lp, lc = pv.lincheck_prove(ch, r1cs, stripe, x_ab)
    och = OracleFs(orc, b"lc-test")
    vp = {"round1_ab": zp["round1_ab"], "round1_c": zp["round1_c"], "rounds": zp["multilinear_rounds"],
          "final_a": zp["final_a_eval"], "final_b": zp["final_b_eval"], "final_c": zp["final_c_eval"]}
    vz = fv.zerocheck_verify(orc, m, vp, och)
    xab = {"z_skip": vz["z"], "x_inner_rest": vz["mlv"][:8], "x_outer": vz["mlv"][8:]}
    vl = fv.lincheck_verify(orc, m, xab, vz["a_eval"], vz["b_eval"], lp, och)
    assert vl["w"] == lc["w"] and vl["r_inner_skip"] == lc["r_inner_skip"] and vl["r_inner_rest"] == lc["r_inner_rest"]
    # w is the witness MLE at (r_inner_skip, r_inner_rest, x_outer)
    zbits = np.frombuffer(z.cpu().numpy().tobytes(), np.uint8)
    mlv = orc.zc_fold_at_z(zbits, m, 6, orc.zc_lagrange_weights(6, lc["r_inner_skip"]))
    assert f128_to_ints(orc.partial_eval_lsb(mlv, lc["r_inner_rest"] + xab["x_outer"]))[0] == lc["w"]


def test_proof_bound_and_small_buffer(orc):
    """flk_blake3_proof_bound bounds the proof of each profile at m = 22 (the slim
    profile commits at rate 1/4); the slim and secure proofs are accepted by the
    restated verifier and equal the oracle prover's; a buffer one byte short of the
    proof is refused."""
    import oracle_prover as op
    comps, blocks = _blocks(8, 3)
    s = setup_for(256)
    proof, _ = s.prove(blocks)
    with pytest.raises(pv.nat.FlockGpuError):
        s.prove(blocks, cap=len(proof) - 1)
    for prof in ("slim", "secure"):
        print(f"case profile={prof}", flush=True)
        t = pv.Blake3Setup(256, profile=prof)
        p2, _ = t.prove(blocks)
        assert len(p2) <= t._cap
        fv.verify_r1cs(orc, fv.parse_bundle(p2))
        want, _ = op.prove(orc, comps.view(COMPRESSION_DT).reshape(-1), 8, threads=8,
                           profile=pv.PROFILES[prof])
        assert p2 == want


@pytest.mark.parametrize("skip", [0, 3, 40, 120])
def test_device_failure_inside_prove_latches(skip):
    """A device failure at any point of a prove (injected after `skip` entry-point
    calls, from the host) makes the prove return FLK_E_DEVICE and latches the
    library; after flk_undisable the same setup proves the same bytes again."""
    lib = pv.nat.load()
    _, blocks = _blocks(8, 51)
    s = setup_for(256)
    want, _ = s.prove(blocks)
    lib.flk_inject_device_failures(skip, 1)
    try:
        with pytest.raises(pv.nat.FlockGpuError) as ei:
            s.prove(blocks)
        assert "status 2" in str(ei.value)  # FLK_E_DEVICE
        assert lib.flk_is_disabled() == 1
        with pytest.raises(pv.nat.FlockGpuError):
            s.prove(blocks)  # FLK_E_DISABLED while latched
    finally:
        lib.flk_inject_device_failures(0, 0)
        lib.flk_undisable()
    assert s.prove(blocks)[0] == want
    assert s.arena_stats()["fallbacks"] == 0


def test_setup_rejects_unsupported_sizes():
    with pytest.raises(pv.nat.FlockGpuError):
        pv.Blake3Setup(1 << 6)  # m = 20: no Ligerito ladder below m = 22
    s = setup_for(256)
    _, blocks = _blocks(9, 1)
    with pytest.raises(ValueError):
        s.prove(blocks)
