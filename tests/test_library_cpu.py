"""CPU-side checks of the device library: it loads, exports every symbol the
header declares, rejects bad shapes without latching, honours the disable
switch, and its host-side tables (twiddles, scalar product) equal the oracle.
No device computation runs here.
"""
import ctypes
import os
import re

import pytest

import flock_challenge_multi_amd._native as nat
import pyfield as pf
from oracle_ffi import f128_to_ints

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "flock_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(flk_[a-z0-9_]+)\s*\(", text)))


def test_exports_every_declared_symbol():
    lib = nat.load()
    names = _declared_symbols()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
        assert n in nat.SIGNATURES, f"{n} has no ctypes signature"


def test_twiddles_match_oracle(orc):
    for dim in [1, 5, 11, 14]:
        buf = (nat.F128 * ((1 << dim) - 1))()
        nat.call("flk_ntt_twiddles", dim, buf)
        assert [buf[i].to_int() for i in range(len(buf))] == f128_to_ints(orc.twiddles(dim))


def test_host_scalar_product():
    import random
    rng = random.Random(5)
    for _ in range(50):
        a, b = rng.getrandbits(128), rng.getrandbits(128)
        assert nat.f128_mul(a, b) == pf.mul(a, b)


def test_disable_switch_and_shape_errors():
    lib = nat.load()
    lib.flk_set_enabled(0)
    try:
        assert lib.flk_is_disabled() == 1
        st = lib.flk_sc_glue(None, None, 4, nat.F128(0, 0), None)
        assert st == nat.FLK_E_DISABLED
        assert b"flk_set_enabled" in lib.flk_last_error()
    finally:
        lib.flk_set_enabled(1)
    assert lib.flk_is_disabled() == 0
    # null pointers / bad shapes are refused before any device work, without latching
    assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_SHAPE
    assert lib.flk_ntt_rs_encode_interleaved(4, None, 8, None, 24, 2, None) == nat.FLK_E_SHAPE
    assert lib.flk_merkle_tree(None, 64, 4, nat.FLK_HASH_SHA256, None, None) in (nat.FLK_E_SHAPE,)
    assert lib.flk_is_disabled() == 0


def test_env_kill_switch(monkeypatch):
    lib = nat.load()
    monkeypatch.setenv("FLOCK_NO_GPU", "1")
    assert lib.flk_is_disabled() == 1
    monkeypatch.setenv("FLOCK_NO_GPU", "0")
    assert lib.flk_is_disabled() == 0


def test_injected_device_failure_latches():
    """A device runtime failure latches the library off for every later call until
    flk_undisable() (the permanent fallback latch of CF/gpu.rs:455-516), forced from
    the host with flk_inject_device_failures: no device is touched."""
    lib = nat.load()
    assert lib.flk_is_disabled() == 0
    lib.flk_inject_device_failures(0, 1)
    try:
        assert lib.flk_is_disabled() == 0  # asking does not consume the injection
        assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_DEVICE
        assert b"injected" in lib.flk_last_error()
        assert lib.flk_is_disabled() == 1
        # latched: every entry point refuses, shape errors included, until undisable
        assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_DISABLED
        assert lib.flk_zc_round1(None, None, None, 20, None, None, None, None) == nat.FLK_E_DISABLED
        assert b"latched" in lib.flk_last_error()
    finally:
        lib.flk_inject_device_failures(0, 0)
        lib.flk_undisable()
    assert lib.flk_is_disabled() == 0
    assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_SHAPE  # back to ordinary refusals
    lib.flk_inject_device_failures(1, 2)  # one call passes, then two fail (the second after undisable)
    try:
        assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_SHAPE
        assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_DEVICE
        lib.flk_undisable()
        assert lib.flk_sc_round_msg(None, None, 8, None, None) == nat.FLK_E_DEVICE
    finally:
        lib.flk_inject_device_failures(0, 0)
        lib.flk_undisable()
    assert lib.flk_is_disabled() == 0


def test_ntt_schedule_registry_and_candidates():
    """The autotuner's candidate schedules are valid groupings of the layers, and
    flk_ntt_set_schedule / flk_ntt_get_schedule install, report and remove one
    (no device work: the registry is host state); malformed schedules are refused
    without latching."""
    import ctypes

    from flock_challenge_multi_amd import _native as nat
    from flock_challenge_multi_amd import autotune
    lib = nat.load()
    for n in range(1, 24):
        cands = autotune.candidates(n)
        assert cands, n
        for c in cands:
            assert sum(c) == n and all(1 <= k <= 3 for k in c), (n, c)
    default = autotune.current(19)
    assert default == (2, 2, 3, 3, 3, 3, 3)
    autotune.install(19, (3, 3, 3, 3, 3, 2, 2))
    assert autotune.current(19) == (3, 3, 3, 3, 3, 2, 2)
    assert autotune.current(19, 3) == default  # another lane count keeps the default
    autotune.install(19, ())
    assert autotune.current(19) == default
    bad = (ctypes.c_int * 3)(3, 3, 3)
    assert lib.flk_ntt_set_schedule(10, 6, bad, 3) == nat.FLK_E_SHAPE  # sums to 9
    five = (ctypes.c_int * 2)(5, 5)
    assert lib.flk_ntt_set_schedule(10, 6, five, 2) == nat.FLK_E_SHAPE  # 5-layer passes
    assert lib.flk_ntt_set_schedule(0, 6, bad, 0) == nat.FLK_E_SHAPE
    assert lib.flk_is_disabled() == 0
