"""Kernel variants behind the library's environment switches (read once per
process) against the oracle: each setting runs tests/variant_check.py in a
child process of its own, one after another."""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
CHECK = os.path.join(HERE, "variant_check.py")

VARIANTS = [
    {"FLOCK_NTT_POS_TABLES": "0"},                                  # deepest pass on nibble-pair tables
    {"FLOCK_NTT_POS_TABLES": "0", "FLOCK_NTT_PAIR_TABLES": "0"},    # ... on 16-entry tables, 4-bit Horner
    {"FLOCK_NTT_SETS": "1"},                                        # one set per thread in shared passes
    {"FLOCK_NTT_SETS": "4"},
    {"FLOCK_NTT_COMB": "0"},                                        # byte comb, two-input XORs
    {"FLOCK_NTT_COMB": "1"},                                        # byte comb, three-input XORs
    {"FLOCK_NTT_SHARED_TABLES": "0"},                               # every pass on per-group tables
    {"FLOCK_NTT_SCHEDULE": "3,3,3,2,2"},                            # 2-layer deepest pass (the 2^13 x 64 commit)
    {"FLOCK_NTT_RADIX_LOG": "2"},
]


@pytest.mark.parametrize("env", VARIANTS, ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()))
def test_ntt_variant_matches_oracle(dev, env):
    print(f"variant {env}", flush=True)
    r = subprocess.run([sys.executable, CHECK, "ntt"], env={**os.environ, **env}, capture_output=True, text=True,
                       timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "MISMATCH" not in r.stdout


# Work-group counts of the product-sum kernels (their sums go through per-work-group
# partials): tiny and uneven grids make every work-group stride over many items.
REDUCTION_VARIANTS = [
    {"FLOCK_SC_GRID": "3", "FLOCK_RD_GRID": "5"},
    {"FLOCK_SC_GRID": "1", "FLOCK_RD_GRID": "1"},
    {"FLOCK_SC_GRID": "65535", "FLOCK_RD_GRID": "65535"},
]


@pytest.mark.parametrize("env", REDUCTION_VARIANTS, ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()))
def test_reduction_variant_matches_oracle(dev, env):
    print(f"variant {env}", flush=True)
    r = subprocess.run([sys.executable, CHECK, "reductions"], env={**os.environ, **env}, capture_output=True,
                       text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "MISMATCH" not in r.stdout


# Word-range parts of the witness kernel (default: 2 at 2^15 compressions, 1 at 2^18).
WITNESS_VARIANTS = [{"FLOCK_WIT_PARTS": v} for v in ("1", "2", "4", "16")]


@pytest.mark.parametrize("env", WITNESS_VARIANTS, ids=lambda e: ",".join(f"{k}={v}" for k, v in e.items()))
def test_witness_variant_matches_oracle(dev, env):
    print(f"variant {env}", flush=True)
    r = subprocess.run([sys.executable, CHECK, "witness"], env={**os.environ, **env}, capture_output=True,
                       text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "MISMATCH" not in r.stdout

