"""GPU test of the allocation-free round message (flk_sc_round_msg_noalloc):
through the C ABI, two calls in a row (the second on a smaller input, so it
reads a shorter prefix of the shared device scratch the first one filled) and
a single-work-group call, each bit-exact against flk_sc_round_msg on the same
device inputs."""
import ctypes

import pytest
import torch

from flock_challenge_multi_amd import _native as nat
from flock_challenge_multi_amd import hotpath as hp

pytestmark = pytest.mark.gpu


def _rand(n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randint(-(2**63), 2**63 - 1, (n, 2), dtype=torch.int64, device="cuda", generator=g)


def _msg(name, f, b):
    out = torch.full((2, 2), -1, dtype=torch.int64, device="cuda")
    nat.call(name, ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(b.data_ptr()), f.shape[0],
             ctypes.c_void_p(out.data_ptr()), hp._stream())
    torch.cuda.synchronize()
    return out.cpu()


def test_noalloc_round_msg_matches_round_msg():
    torch.cuda.set_device(0)
    # 2^23 elements: 2^22 pairs, the grid capped at 4096 work-groups (the whole scratch)
    # then 2^15 elements: 128 work-groups, a prefix of the scratch the first call filled
    # then 2 elements: one work-group, which writes its output directly
    for n, seed in ((1 << 23, 1), (1 << 15, 2), (2, 3)):
        print(f"case noalloc round_msg n={n}", flush=True)
        f, b = _rand(n, seed), _rand(n, seed + 100)
        want = _msg("flk_sc_round_msg", f, b)
        got = _msg("flk_sc_round_msg_noalloc", f, b)
        assert torch.equal(got, want)


def test_noalloc_rejects_oversize():
    f = _rand(4, 9)
    with pytest.raises(nat.FlockGpuError):
        nat.call("flk_sc_round_msg_noalloc", ctypes.c_void_p(f.data_ptr()), ctypes.c_void_p(f.data_ptr()),
                 (1 << 40) * 2, ctypes.c_void_p(f.data_ptr()), hp._stream())
