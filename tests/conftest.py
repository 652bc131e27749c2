import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a real MI355X (runs the HIP library)")


def pytest_collection_modifyitems(config, items):
    # A `-m gpu` run on a machine without a GPU should fail loudly rather than
    # pass silently, so nothing is auto-skipped here.
    pass


@pytest.fixture(scope="session")
def orc():
    import oracle_ffi
    return oracle_ffi.Oracle()


@pytest.fixture(scope="session")
def dev():
    """The device library on cuda:0 (GPU tests only)."""
    import torch
    assert torch.cuda.is_available(), "gpu tests need a HIP device"
    import flock_challenge_multi_amd._native as nat
    lib = nat.load()
    assert lib.flk_device_available() == 1, nat.load().flk_last_error()
    return lib
