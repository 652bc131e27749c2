"""Runs the C++ host-mirror parity program (tests/cpp/host_mirror_test.cpp,
built by __graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "host_mirror_test")


def test_cpp_host_mirror_parity(dev):
    assert os.path.exists(EXE), "build it with __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=110)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "host mirror parity ok" in r.stdout
