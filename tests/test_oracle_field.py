"""Pins the oracle's GF(2^128) / GF(2^8) / phi8 arithmetic.

Checks: the reference's literal KATs (CF/field/gf2_128.rs:468-494), an
independent long-division implementation (tests/pyfield.py), field axioms,
the inverse, the PCLMUL path against the bit-serial reference form, and phi8
against the reference table fixture plus its defining properties.
"""
import json
import os
import random

import pyfield as pf

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def test_ghash_reduction_smoking_gun(orc):
    # CF/field/gf2_128.rs:468-494
    x = 2
    assert orc.mul(x, 1 << 127) == 0x87
    assert orc.mul(x, 1 << 63) == 1 << 64
    assert orc.mul(1 << 64, 1 << 64) == 0x87
    assert orc.mul(x, x) == 4


def test_mul_matches_long_division(orc):
    rng = random.Random(1)
    for _ in range(300):
        a, b = rng.getrandbits(128), rng.getrandbits(128)
        assert orc.mul(a, b) == pf.mul(a, b)
    # high-bit inputs (every reduction path)
    for a in [(1 << 128) - 1, 1 << 127, (1 << 127) | 1, 0xFF << 120]:
        for b in [(1 << 128) - 1, 1 << 127, 3 << 126]:
            assert orc.mul(a, b) == pf.mul(a, b)


def test_pclmul_matches_bitwise(orc):
    rng = random.Random(2)
    for _ in range(500):
        a, b = rng.getrandbits(64), rng.getrandbits(64)
        assert orc.clmul64(a, b) == orc.clmul64(a, b, bitwise=True) == pf.clmul(a, b)


def test_field_axioms_and_inverse(orc):
    rng = random.Random(7)
    for _ in range(64):
        a, b, c = (rng.getrandbits(128) for _ in range(3))
        assert orc.mul(orc.mul(a, b), c) == orc.mul(a, orc.mul(b, c))
        assert orc.mul(a, b ^ c) == orc.mul(a, b) ^ orc.mul(a, c)
        assert orc.mul(a, b) == orc.mul(b, a)
        if a:
            assert orc.mul(a, orc.inv(a)) == 1
    assert orc.inv(0) == 0
    # Fermat: a^(2^128) = a
    a = rng.getrandbits(128)
    assert pf.pow_(a, 1 << 128) == a


def test_mul_by_x(orc):
    rng = random.Random(3)
    for _ in range(100):
        a = rng.getrandbits(128)
        assert orc.mul_by_x(a) == orc.mul(a, 2)


def test_f8_matches_long_division(orc):
    for a in range(256):
        for b in range(0, 256, 7):
            assert orc.f8_mul(a, b) == pf.mul8(a, b)
    # FIPS-197 worked example: {57} * {83} = {c1}
    assert orc.f8_mul(0x57, 0x83) == 0xC1


def test_phi8_table_matches_reference_fixture(orc):
    t = orc.phi8_table()
    fix = json.load(open(os.path.join(GOLDEN, "phi8_table.json")))["table"]
    for v in range(256):
        lo, hi = int(fix[v][0], 16), int(fix[v][1], 16)
        assert (int(t[v]["lo"]), int(t[v]["hi"])) == (lo, hi), v


def test_phi8_is_a_field_embedding(orc):
    t = orc.phi8_table()
    phi = [int(e["lo"]) | (int(e["hi"]) << 64) for e in t]
    assert phi[0] == 0 and phi[1] == 1
    g = phi[2]
    # phi8(x) is a root of the AES polynomial x^8 + x^4 + x^3 + x + 1 in GF(2^128)
    p = [pf.pow_(g, k) for k in range(9)]
    assert p[8] ^ p[4] ^ p[3] ^ p[1] ^ 1 == 0
    rng = random.Random(4)
    for _ in range(400):
        a, b = rng.randrange(256), rng.randrange(256)
        assert phi[a ^ b] == phi[a] ^ phi[b]
        assert phi[orc.f8_mul(a, b)] == pf.mul(phi[a], phi[b])
