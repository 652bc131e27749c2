"""GPU parity: proof-of-work grinding (flk_pow_grind, flk_pow_scan) against
the oracle through the C ABI: the same smallest nonce for many transcript
digests and difficulties, range scans with offsets, empty and saturated ranges
and the zero-bit cases, and shape errors."""
import random
import sys

import pytest

from flock_challenge_multi_amd import _native as nat
from flock_challenge_multi_amd import hotpath as hp

pytestmark = pytest.mark.gpu


def _digest(seed):
    r = random.Random(seed)
    return bytes(r.randrange(256) for _ in range(32))


@pytest.mark.parametrize("bits", [1, 2, 5, 8, 13, 16, 20, 22])
def test_grind_matches_oracle(orc, dev, bits):
    print(f"pow_grind bits={bits}", flush=True)
    for seed in range(6 if bits <= 16 else 2):
        d = _digest(1000 * bits + seed)
        assert hp.grind_pow_nonce(d, bits) == orc.pow_grind(d, bits), seed


def test_grind_zero_bits_and_limits(orc, dev):
    d = _digest(5)
    assert hp.grind_pow_nonce(d, 0) == 0
    import ctypes
    out = ctypes.c_uint64(7)
    buf = (ctypes.c_uint8 * 32).from_buffer_copy(d)
    assert dev.flk_pow_grind(buf, 49, ctypes.byref(out), None) == nat.FLK_E_SHAPE
    assert dev.flk_pow_scan(buf, 0, 10, 65, ctypes.byref(out), None) == nat.FLK_E_SHAPE
    assert dev.flk_pow_grind(None, 5, ctypes.byref(out), None) == nat.FLK_E_SHAPE
    assert dev.flk_is_disabled() == 0


def test_grind_large_is_minimal(orc, dev):
    """At 26 bits the oracle only checks the answer; the device scan of every
    smaller nonce finds nothing."""
    d = _digest(77)
    n = hp.grind_pow_nonce(d, 26)
    assert orc.pow_ok(d, n, 26)
    assert n == 0 or hp.pow_scan(d, 0, n, 26) is None


def test_scan_ranges(orc, dev):
    d = _digest(9)
    bits = 10
    first = orc.pow_grind(d, bits)
    # the smallest in a range that starts past the first match
    second = hp.pow_scan(d, first + 1, 1 << 20, bits)
    assert second is not None and second > first and orc.pow_ok(d, second, bits)
    for m in range(first + 1, second):
        assert not orc.pow_ok(d, m, bits)
    assert hp.pow_scan(d, first, 1, bits) == first
    assert hp.pow_scan(d, 0, first, bits) is None
    assert hp.pow_scan(d, 123, 0, bits) is None            # empty range
    assert hp.pow_scan(d, 123, 5, 0) == 123                 # every nonce passes at 0 bits
    assert hp.pow_scan(d, 0, 1 << 16, 64) is None          # no 64-bit match in 2^16 nonces
    top = 2**64 - 1000
    assert hp.pow_scan(d, top, 10**6, 0) == top             # the range is capped at 2^64 - 1
    got = hp.pow_scan(d, top, 10**6, 4)
    want = next((n for n in range(top, 2**64 - 1) if orc.pow_ok(d, n, 4)), None)
    assert got == want


def test_scan_many_offsets(orc, dev):
    r = random.Random(3)
    for _ in range(20):
        d = _digest(r.randrange(10**6))
        start = r.randrange(2**40)
        bits = r.randrange(1, 9)
        want = next((n for n in range(start, start + 3000) if orc.pow_ok(d, n, bits)), None)
        assert hp.pow_scan(d, start, 3000, bits) == want


@pytest.mark.parametrize("bits", [0, 5, 10, 14, 18])
def test_transcript_grind_roundtrip(orc, dev, bits):
    """fs_challenger_pow_roundtrip (CF/challenger.rs:2185-2209) with the device
    grind: the verifier accepts the nonce, the nonce is the oracle's, and the
    challenges after it agree."""
    def script():
        t = hp.FsChallenger(b"pow-test")
        t.observe_label(b"flock-pow-test")
        t.observe_bytes(b"some root data")
        return t
    prover, verifier = script(), script()
    d = prover.state_digest()
    nonce = prover.grind_pow(bits)
    assert nonce == orc.pow_grind(d, bits)
    assert verifier.verify_pow(nonce, bits)
    for _ in range(4):
        assert prover.sample_f128() == verifier.sample_f128()
    if bits:
        wrong = script()
        assert not wrong.verify_pow(nonce + 1, bits) or orc.pow_ok(d, nonce + 1, bits)
