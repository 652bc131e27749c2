"""Writes tests/golden/proof_m22_seed123c.bin: the device prover's proof of
generate_compressions(8, 0x123c) (m = 22) on the transcript domain
b"flock-bench-v0".  Needs a GPU; run from the repository root:
    python tests/golden/make_proof_fixture.py
The CPU tests verify this fixture with the restated verifier and check that
every single-field mutation of it is rejected; the GPU tests check that the
device prover still reproduces it byte for byte."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from flock_challenge_multi_amd import prover as pv  # noqa: E402
from flock_challenge_multi_amd.inputs import generate_compressions  # noqa: E402

LOG2, SEED = 8, 0x123C

if __name__ == "__main__":
    blocks = torch.from_numpy(generate_compressions(LOG2, SEED).view(np.int32)).cuda()
    proof, _ = pv.Blake3Setup(1 << LOG2).prove(blocks)
    path = os.path.join(ROOT, "tests", "golden", "proof_m22_seed123c.bin")
    open(path, "wb").write(proof)
    print(path, len(proof), "bytes")
