"""Pins the oracle's BLAKE3 R1CS witness generator (oracle/witness.c).

The reference (CP/r1cs_hashes/blake3.rs, CP = crates/flock-prover/src/) fixes
the witness by four things, each checked here against something other than
the oracle's own formulas:

* the layout constants its own test asserts (blake3.rs:5596-5610);
* the compression itself: out_lo of a root-flagged single-block compression is
  the BLAKE3 hash, pinned by the published BLAKE3 vectors (the reference's
  compress_matches_blake3_crate_* tests, :5615-5652), and out_lo / out_hi equal
  the finalisation XORs of the compression (witness_encodes_correct_output,
  :5655-5680);
* a = A z and b = B z for the matrices of build_matrices (:483-660), rebuilt
  here symbolically (each row a GF(2) combination of witness positions, kept
  as a Python int bitset) — the reference's
  generate_witness_with_ab_packed_matches_apply (:5726-5752) with an
  independent construction of A and B;
* R1CS satisfaction a * b = z (C = I) of honest witnesses and its failure on a
  flipped carry bit (:5683-5723).

The lincheck stripe is checked against a numpy bit transpose of z, and the
padding slots against the all-zero compression's witness.
"""
import json
import os

import numpy as np
import pytest

from oracle_ffi import COMPRESSION_DT

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")

K = 1 << 14
CONST = 512
GS_BASE = 1153
G_STRIDE = 250
OUT_HI_BASE = 15153
USEFUL = 15409
IV = [0x6A09E667, 0xBB67AE85, 0x3C6EF372, 0xA54FF53A, 0x510E527F, 0x9B05688C, 0x1F83D9AB, 0x5BE0CD19]
PERM = [2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8]
LANES = [(0, 4, 8, 12), (1, 5, 9, 13), (2, 6, 10, 14), (3, 7, 11, 15),
         (0, 5, 10, 15), (1, 6, 11, 12), (2, 7, 8, 13), (3, 4, 9, 14)]
CHUNK_START, CHUNK_END, ROOT = 1, 2, 8


def random_compressions(seed, n):
    rng = np.random.default_rng(seed)
    c = np.zeros(n, COMPRESSION_DT)
    c["cv"] = rng.integers(0, 2**32, (n, 8), dtype=np.uint32)
    c["m"] = rng.integers(0, 2**32, (n, 16), dtype=np.uint32)
    c["counter"] = rng.integers(0, 2**63, n, dtype=np.uint64) * 2 + rng.integers(0, 2, n, dtype=np.uint64)
    c["block_len"] = rng.integers(0, 2**32, n, dtype=np.uint32)
    c["flags"] = rng.integers(0, 2**32, n, dtype=np.uint32)
    return c


# ---------------------------------------------------------------------------
# Symbolic A / B (build_matrices restated over bitsets).
# ---------------------------------------------------------------------------

def _slot(base):
    return [1 << (base + i) for i in range(32)]


def _const(c):
    return [(1 << CONST) if (c >> i) & 1 else 0 for i in range(32)]


def _xor(x, y):
    return [p ^ q for p, q in zip(x, y)]


def _rotr(x, n):
    return [x[(i + n) % 32] for i in range(32)]


def build_matrices():
    """Rows of A and B as bitsets over the K witness positions of one block."""
    A = [0] * K
    B = [0] * K
    A[CONST] = B[CONST] = 1 << CONST
    inputs = [(0, 256), (513, 512), (1025, 32), (1057, 32), (1089, 32), (1121, 32)]
    for base, n in inputs:
        for s in range(base, base + n):
            A[s] = 1 << s
            B[s] = 1 << CONST

    def add(x, y, cb):
        cin = 0
        out = []
        for i in range(32):
            if i < 31:
                A[cb + i] = x[i] ^ cin
                B[cb + i] = y[i] ^ cin
            out.append(x[i] ^ y[i] ^ cin)
            if i < 31:
                cin ^= 1 << (cb + i)
        return out

    def lin(s0, word):
        for i in range(32):
            A[s0 + i] = word[i]
            B[s0 + i] = 1 << CONST

    st = [_slot(32 * w) for w in range(8)] + [_const(IV[i]) for i in range(4)]
    st += [_slot(1025), _slot(1057), _slot(1089), _slot(1121)]
    perm = list(range(16))
    for r in range(7):
        for gi in range(8):
            g = 8 * r + gi
            base = GS_BASE + G_STRIDE * g
            la, lb, lc, ld = LANES[gi]
            mx, my = _slot(513 + 32 * perm[2 * gi]), _slot(513 + 32 * perm[2 * gi + 1])
            a, b, c, d = st[la], st[lb], st[lc], st[ld]
            tmp0 = add(a, b, base)
            a1 = add(tmp0, mx, base + 31)
            d1 = _rotr(_xor(d, a1), 16)
            c1 = add(c, d1, base + 62)
            b1 = _rotr(_xor(b, c1), 12)
            tmp1 = add(a1, b1, base + 93)
            a2 = add(tmp1, my, base + 124)
            d2 = _rotr(_xor(d1, a2), 8)
            c2 = add(c1, d2, base + 155)
            lin(base + 186, _rotr(_xor(b1, c2), 7))
            lin(base + 218, d2)
            st[la], st[lb], st[lc], st[ld] = a2, _slot(base + 186), c2, _slot(base + 218)
        perm = [perm[PERM[i]] for i in range(16)]
    for w in range(8):
        lin(256 + 32 * w, _xor(st[w], st[w + 8]))
        lin(OUT_HI_BASE + 32 * w, _xor(st[w + 8], _slot(32 * w)))
    return A, B


@pytest.fixture(scope="module")
def matrices():
    return build_matrices()


def _as_int(words):
    return int.from_bytes(np.asarray(words, np.uint64).tobytes(), "little")


def _apply(rows, zint):
    v = 0
    for s, row in enumerate(rows):
        if row and (row & zint).bit_count() & 1:
            v |= 1 << s
    return v


# ---------------------------------------------------------------------------

def test_layout_constants(orc):
    assert orc.witness_layout() == [K, CONST, GS_BASE, G_STRIDE, OUT_HI_BASE, USEFUL]
    assert OUT_HI_BASE == GS_BASE + 56 * G_STRIDE and USEFUL == OUT_HI_BASE + 256 <= K


def _word(words, pos):
    zint = _as_int(words)
    return (zint >> pos) & 0xFFFFFFFF


def test_out_lo_is_the_published_hash(orc):
    v = json.load(open(os.path.join(GOLDEN, "blake3_vectors.json")))
    cases = [(bytes(i % 251 for i in range(int(n))), h) for n, h in v["pattern_inputs"].items()
             if int(n) <= 64]
    cases += [(s.encode(), h) for s, h in v["plain_inputs"].items()]
    assert len(cases) >= 3
    for data, h in cases:
        block = data + bytes(64 - len(data))
        c = np.zeros(1, COMPRESSION_DT)
        c["cv"] = IV
        c["m"] = np.frombuffer(block, "<u4")
        c["block_len"] = len(data)
        c["flags"] = CHUNK_START | CHUNK_END | ROOT
        z, a, b = orc.block_witness(c)
        got = b"".join(_word(z, 256 + 32 * w).to_bytes(4, "little") for w in range(8))
        assert got.hex() == h


def test_outputs_are_the_compression_xors(orc):
    comps = random_compressions(11, 6)
    for c in comps:
        z, _, _ = orc.block_witness(c)
        block = np.asarray(c["m"], "<u4").tobytes()
        out = orc.compress(c["cv"], block, int(c["counter"]), int(c["block_len"]), int(c["flags"]))
        for w in range(8):
            assert _word(z, 256 + 32 * w) == out[w]
            assert _word(z, OUT_HI_BASE + 32 * w) == out[w + 8]


def test_inputs_are_placed(orc):
    c = random_compressions(12, 1)[0]
    z, a, b = orc.block_witness(c)
    for w in range(8):
        assert _word(z, 32 * w) == c["cv"][w]
    for i in range(16):
        assert _word(z, 513 + 32 * i) == c["m"][i]
    assert _word(z, 1025) == int(c["counter"]) & 0xFFFFFFFF
    assert _word(z, 1057) == int(c["counter"]) >> 32
    assert _word(z, 1089) == c["block_len"] and _word(z, 1121) == c["flags"]
    assert (_as_int(z) >> CONST) & 1 == 1


@pytest.mark.parametrize("seed", [0, 1])
def test_ab_equal_matrix_products(orc, matrices, seed):
    A, B = matrices
    comps = random_compressions(100 + seed, 2)
    for c in comps:
        z, a, b = orc.block_witness(c)
        zi = _as_int(z)
        assert _apply(A, zi) == _as_int(a)
        assert _apply(B, zi) == _as_int(b)


def test_zero_compression_ab(orc, matrices):
    A, B = matrices
    z, a, b = orc.block_witness(np.zeros(1, COMPRESSION_DT))
    zi = _as_int(z)
    assert _apply(A, zi) == _as_int(a) and _apply(B, zi) == _as_int(b)


def test_honest_witness_satisfies_r1cs(orc):
    comps = random_compressions(21, 5)
    z, a, b, _ = orc.witness(comps, 3, stripe=False)
    assert np.array_equal(a & b, z)
    # padding rows are all zero in every slot
    zb = np.unpackbits(z.view(np.uint8).reshape(8, -1), axis=1, bitorder="little")
    assert not zb[:, USEFUL:].any()
    assert zb[:, CONST].all()


def test_flipped_carry_violates_r1cs(orc, matrices):
    A, B = matrices
    c = random_compressions(22, 1)[0]
    z, _, _ = orc.block_witness(c)
    zi = _as_int(z)
    assert _apply(A, zi) & _apply(B, zi) == zi
    pos = GS_BASE + G_STRIDE * 10 + 31 * 4 + 5  # G 10, add a2, bit 5
    zf = zi ^ (1 << pos)
    assert _apply(A, zf) & _apply(B, zf) != zf


def test_padding_slots_and_stripe(orc):
    comps = random_compressions(31, 13)
    z, a, b, st = orc.witness(comps, 4)
    zero = orc.block_witness(np.zeros(1, COMPRESSION_DT))
    for j in range(16):
        ref = orc.block_witness(comps[j]) if j < 13 else zero
        for got, want in zip((z, a, b), ref):
            assert np.array_equal(got[256 * j: 256 * (j + 1)], want)
    bits = np.unpackbits(z.view(np.uint8).reshape(16, -1), axis=1, bitorder="little")
    want = np.packbits(bits.reshape(2, 8, K).transpose(0, 2, 1), axis=-1, bitorder="little")
    assert np.array_equal(st, want.reshape(-1))


def test_shapes_rejected(orc):
    comps = random_compressions(1, 9)
    with pytest.raises(ValueError):
        orc.witness(comps[:4], 2)
    with pytest.raises(ValueError):
        orc.witness(comps, 3)
    z, a, b, st = orc.witness(comps[:0], 3)  # no compressions: all padding
    zero = orc.block_witness(np.zeros(1, COMPRESSION_DT))
    assert np.array_equal(z[:256], zero[0])


def test_word_form_equals_bit_form(orc):
    """The two restatements (bit by bit; word records as the reference's packed
    builder) agree on whole batches, stripe included, on 1 and 5 threads."""
    comps = random_compressions(41, 37)
    want = orc.witness(comps, 6)
    for threads in (1, 5):
        got = orc.witness_mt(comps, 6, threads)
        for g, w in zip(got, want):
            assert np.array_equal(g, w)
    zero = np.zeros(1, COMPRESSION_DT)
    z2, a2, b2 = (np.zeros(256, np.uint64) for _ in range(3))
    orc.L.orc_blake3_block_witness_words(zero.ctypes.data, z2.ctypes.data, a2.ctypes.data, b2.ctypes.data)
    for g, w in zip((z2, a2, b2), orc.block_witness(zero)):
        assert np.array_equal(g, w)
