"""The host Fiat-Shamir transcript (flk_fs_*, csrc/transcript.cpp) against a
byte-stream model of FsChallenger over BLAKE3 (CF/challenger.rs:170-177,
226-383,502-523): the model keeps every absorbed byte, and its digests and
challenges come from the oracle's BLAKE3 hash and XOF (pinned by the published
BLAKE3 vectors), so the library's incremental hasher (CV stack, chunk and
block boundaries, output blocks past 64 bytes) is checked against the
oracle's recursive tree.  Host code only: runs without a GPU."""
import random
import struct

import pytest

from flock_challenge_multi_amd import hotpath as hp


class Model:
    def __init__(self, orc, domain):
        self.orc = orc
        self.buf = b"\x01" + struct.pack("<Q", len(domain)) + domain

    def label(self, l):
        self.buf += b"\x02" + struct.pack("<Q", len(l)) + l

    def f128(self, v):
        self.buf += b"\x03\x01" + struct.pack("<QQ", v & (2**64 - 1), v >> 64)

    def slice(self, vs):
        self.buf += b"\x03\x02" + struct.pack("<Q", len(vs))
        for v in vs:
            self.buf += struct.pack("<QQ", v & (2**64 - 1), v >> 64)

    def bytes_(self, b):
        self.buf += b"\x05" + struct.pack("<Q", len(b)) + b

    def _squeeze(self, n):
        out = self.orc.blake3_xof(self.buf, 16 * n)
        self.buf += out
        return [int.from_bytes(out[16 * i:16 * i + 16], "little") for i in range(n)]

    def sample(self):
        self.buf += b"\x04\x01"
        return self._squeeze(1)[0]

    def sample_vec(self, n):
        self.buf += b"\x04\x02" + struct.pack("<Q", n)
        return self._squeeze(n)

    def digest(self):
        return self.orc.blake3(self.buf)


def test_fresh_and_empty_domain(orc):
    for dom in (b"", b"d", b"flock-r1cs-v0", bytes(range(200)) * 7):
        assert hp.FsChallenger(dom).state_digest() == Model(orc, dom).digest()


@pytest.mark.parametrize("seed", range(6))
def test_random_scripts(orc, seed):
    r = random.Random(seed)
    dom = bytes(r.randrange(256) for _ in range(r.randrange(40)))
    t, m = hp.FsChallenger(dom), Model(orc, dom)
    for _ in range(60):
        op = r.randrange(7)
        if op == 0:
            l = bytes(r.randrange(256) for _ in range(r.randrange(30)))
            t.observe_label(l), m.label(l)
        elif op == 1:
            v = r.getrandbits(128)
            t.observe_f128(v), m.f128(v)
        elif op == 2:
            vs = [r.getrandbits(128) for _ in range(r.randrange(70))]
            t.observe_f128_slice(vs), m.slice(vs)
        elif op == 3:
            # long payloads cross block and chunk boundaries (1 KiB chunks)
            b = bytes(r.randrange(256) for _ in range(r.choice([0, 1, 63, 64, 65, 1000, 1024, 2500])))
            t.observe_bytes(b), m.bytes_(b)
        elif op == 4:
            assert t.sample_f128() == m.sample()
        elif op == 5:
            n = r.choice([0, 1, 2, 3, 4, 5, 8, 17])
            assert t.sample_f128_vec(n) == m.sample_vec(n)
        else:
            assert t.state_digest() == m.digest()
    assert t.state_digest() == m.digest()
    assert len(m.buf) > 1024  # the scripts reach the multi-chunk tree


def test_clone_diverges_independently(orc):
    t = hp.FsChallenger(b"pow-test")
    t.observe_bytes(b"root")
    c = t.clone()
    a = t.sample_f128()
    assert c.sample_f128() == a
    t.observe_label(b"x")
    assert t.state_digest() != c.state_digest()


def test_verify_pow_on_host(orc):
    for bits in (0, 3, 9):
        t = hp.FsChallenger(b"pow-test")
        t.observe_label(b"flock-pow-test")
        t.observe_bytes(b"some root data")
        d = t.state_digest()
        n = orc.pow_grind(d, bits)
        good, bad = t.clone(), t.clone()
        assert good.verify_pow(n, bits)
        assert not bad.verify_pow(n + 1 if bits else 1, bits) or orc.pow_ok(d, n + 1, bits)
        # the nonce is absorbed either way: the states then differ only by it
        m = Model(orc, b"pow-test")
        m.label(b"flock-pow-test"), m.bytes_(b"some root data"), m.bytes_(struct.pack("<Q", n))
        assert good.state_digest() == m.digest()


def test_sample_distinct_queries(orc):
    for block_len, count in ((1 << 17, 148), (64, 64), (1000, 3), (7, 0)):
        t, m = hp.FsChallenger(b"q"), Model(orc, b"q")
        t.observe_bytes(b"root"), m.bytes_(b"root")
        got = t.sample_distinct_queries(block_len, count)
        want = []
        while len(want) < count:
            q = (m.sample() & (2**64 - 1)) % block_len
            if q not in want:
                want.append(q)
        assert got == sorted(want)
        assert t.state_digest() == m.digest()
