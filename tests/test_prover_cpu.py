"""CPU tests of the proof format and verifier: the statement digest of the
product library against the restated one, the Ligerito ladder data against the
reference's TOML files, and the restated verifier on a device proof fixture
(tests/golden/proof_m22_seed123c.bin, written by tests/golden/make_proof_fixture.py):
reproduced byte for byte by the oracle's CPU prover, accepted as is, rejected after any single-field mutation (the reference's
tamper tests, CF/zerocheck.rs:2096-2170, CF/lincheck.rs:7195-7336,
CF/pcs/ring_switch.rs:5199-5220)."""
import copy
import ctypes
import os

import numpy as np
import pytest

import flock_verifier as fv
from oracle_ffi import COMPRESSION_DT
from flock_challenge_multi_amd.inputs import generate_compressions

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "proof_m22_seed123c.bin")
LOG2, SEED = 8, 0x123C
REF_CFG = "/root/reference/crates/flock-core/configs/ligerito"


@pytest.fixture(scope="module")
def fixture_bytes():
    return open(FIXTURE, "rb").read()


@pytest.fixture(scope="module")
def comps():
    return generate_compressions(LOG2, SEED).view(COMPRESSION_DT).reshape(-1)


@pytest.mark.parametrize("m", [17, 22, 26, 32])
def test_statement_digest_matches_restatement(orc, m):
    from flock_challenge_multi_amd import _native as nat
    out = (ctypes.c_uint8 * 32)()
    nat.call("flk_blake3_statement_digest", m, out)
    assert bytes(out) == fv.statement_digest(orc, m)


def test_statement_digest_depends_on_m(orc):
    assert fv.statement_digest(orc, 22) != fv.statement_digest(orc, 23)


def test_matrix_nonzero_counts():
    a_ptr, a_cols, b_ptr, b_cols = fv.matrices()
    # "BLAKE3's R1CS has ~21M nonzeros" (CF/r1cs.rs:330-333)
    assert 20_000_000 < len(a_cols) + len(b_cols) < 22_000_000
    for r in range(0, 16384, 97):
        assert np.all(np.diff(a_cols[a_ptr[r]:a_ptr[r + 1]].astype(np.int64)) > 0)


@pytest.mark.skipif(not os.path.isdir(REF_CFG), reason="the reference's TOML ladders are not present here")
def test_ladders_match_reference_toml():
    import sys
    sys.path.insert(0, os.path.join(fv.ROOT, "flock_challenge_multi_amd"))
    import gen_ligerito_ladders as g
    lad = fv.ladders()
    for m in range(22, 36):
        for p, prof in enumerate(("fast", "slim", "secure")):
            top, levels, final = g.parse(open(os.path.join(REF_CFG, f"m{m}_{prof}.toml")).read())
            cfg = lad[(m, p)]
            assert cfg["initial_k"] == top["initial_k"] and cfg["yr_log_n"] == final["yr_log_n"]
            assert cfg["queries"] == [l["queries"] for l in levels]
            assert cfg["recursive_ks"] == [l["k_recursive"] for l in levels[1:]]
            assert cfg["fold_grinding_bits"] == [l["fold_grinding_bits"] for l in levels]
            assert cfg["ood_samples"] == [l["ood_samples"] for l in levels]
            assert cfg["log_inv_rates"] == [l["log_inv_rate"] for l in levels]
            assert cfg["grinding_bits"] == [l["grinding_bits"] for l in levels]


def test_oracle_prover_reproduces_fixture(orc, fixture_bytes, comps):
    """The oracle's CPU prover (tests/oracle_prover.py), which shares no code with the
    device library, writes the device fixture byte for byte; threaded round 1 too."""
    import oracle_prover as op
    for threads in (1, 4):
        proof, _ = op.prove(orc, comps, LOG2, threads=threads)
        assert proof == fixture_bytes


def test_fixture_round_trips(fixture_bytes):
    assert fv.serialize_bundle(fv.parse_bundle(fixture_bytes)) == fixture_bytes


def test_fixture_verifies(orc, fixture_bytes, comps):
    assert len(fixture_bytes) < fv.MAX_PROOF_BYTES
    fv.verify_proof(orc, fixture_bytes, comps, LOG2)


def _flip(v):
    return v ^ (1 << 77)


def _set_pair(lst, i, j):
    p = list(lst[i])
    p[j] = _flip(p[j])
    lst[i] = tuple(p)


MUTATIONS = {
    "root": lambda b: b.__setitem__("root", bytes([b["root"][0] ^ 1]) + b["root"][1:]),
    "round1_ab": lambda b: b["zc"]["round1_ab"].__setitem__(3, _flip(b["zc"]["round1_ab"][3])),
    "round1_c": lambda b: b["zc"]["round1_c"].__setitem__(0, _flip(b["zc"]["round1_c"][0])),
    "zc_round_g1": lambda b: _set_pair(b["zc"]["rounds"], 5, 0),
    "zc_round_last_inf": lambda b: _set_pair(b["zc"]["rounds"], -1, 1),
    "final_a": lambda b: b["zc"].__setitem__("final_a", _flip(b["zc"]["final_a"])),
    "final_c": lambda b: b["zc"].__setitem__("final_c", _flip(b["zc"]["final_c"])),
    "lc_round": lambda b: _set_pair(b["lc"]["rounds"], 2, 0),
    "z_partial": lambda b: b["lc"]["z_partial"].__setitem__(7, _flip(b["lc"]["z_partial"][7])),
    "s_hat_v_ab": lambda b: b["s_hat_v"][0].__setitem__(5, _flip(b["s_hat_v"][0][5])),
    "s_hat_v_c": lambda b: b["s_hat_v"][1].__setitem__(100, _flip(b["s_hat_v"][1][100])),
    "l0_row": lambda b: b["lig"]["initial_proof"][0][3].__setitem__(2, (1, 2)),
    "l0_merkle": lambda b: b["lig"]["initial_proof"][1].__setitem__(0, bytes(32)),
    "recursive_root": lambda b: b["lig"]["recursive_roots"].__setitem__(0, bytes(32)),
    "recursive_row": lambda b: b["lig"]["recursive_proofs"][0][0][1].__setitem__(0, (5, 5)),
    "sumcheck_msg": lambda b: _set_pair(b["lig"]["sumcheck_transcript"], 4, 0),
    "fold_nonce": lambda b: b["lig"]["fold_grinding_nonces"].__setitem__(0, b["lig"]["fold_grinding_nonces"][0] + 1),
    "ood_value": lambda b: b["lig"]["ood_values"].__setitem__(0, _flip(b["lig"]["ood_values"][0])),
    "yr": lambda b: b["lig"]["yr"].__setitem__(1, _flip(b["lig"]["yr"][1])),
    "final_row": lambda b: b["lig"]["final_proof"][0][0].__setitem__(0, (9, 9)),
    "final_merkle": lambda b: b["lig"]["final_proof"][1].__setitem__(-1, bytes(32)),
}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_mutation_rejected(orc, fixture_bytes, comps, name):
    b = copy.deepcopy(fv.parse_bundle(fixture_bytes))
    MUTATIONS[name](b)
    mutated = fv.serialize_bundle(b)
    assert mutated != fixture_bytes
    with pytest.raises(fv.Reject):
        fv.verify_proof(orc, mutated, comps, LOG2)


@pytest.mark.parametrize("edit", ["truncate", "trailing", "magic", "version", "flavor"])
def test_malformed_bytes_rejected(fixture_bytes, edit):
    b = bytearray(fixture_bytes)
    if edit == "truncate":
        b = b[:-5]
    elif edit == "trailing":
        b += b"\0"
    elif edit == "magic":
        b[0] ^= 1
    elif edit == "version":
        b[5] = 3
    else:
        b[6] = 3
    with pytest.raises(fv.Reject):
        fv.parse_bundle(bytes(b))


def test_oversize_rejected():
    with pytest.raises(fv.Reject):
        fv.parse_bundle(b"FLOCK\x04\x02" + bytes(fv.MAX_PROOF_BYTES))


def test_wrong_inputs_rejected(orc, fixture_bytes):
    other = generate_compressions(LOG2, SEED + 1).view(COMPRESSION_DT).reshape(-1)
    with pytest.raises(fv.Reject):
        fv.verify_proof(orc, fixture_bytes, other, LOG2)
