"""Host check of the per-compression witness code that the device kernel runs
(csrc/witness_lane.hpp) against the oracle: tests/cpp/witness_lane_test.cpp,
compiled here with g++ (no GPU involved)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_witness_lane_program():
    sys.path.insert(0, ROOT)
    from flock_challenge_multi_amd import build as b
    exe = b.build_witness_lane_test()
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "witness lane ok" in r.stdout
