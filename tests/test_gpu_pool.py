"""GPU tests of batched and served proving (flk_prover_pool_*, csrc/pool.hip): a
proof made in a mix of concurrent jobs of different sizes, domains and profiles
equals, byte for byte, the same proof made alone (flk_blake3_prove); admission
keeps the reserved device memory within the budget; a job larger than the budget
fails alone and the pool keeps serving."""
import numpy as np
import pytest
import torch

from flock_challenge_multi_amd import prover as pv
from flock_challenge_multi_amd.inputs import generate_compressions

pytestmark = pytest.mark.gpu


def _blocks(n, seed):
    log2 = max(3, (n - 1).bit_length())
    c = np.ascontiguousarray(generate_compressions(log2, seed)[:n])
    return torch.from_numpy(c.view(np.int32)).cuda()


def _alone(blocks, domain=pv.DOMAIN, profile="fast"):
    return pv.Blake3Setup(blocks.shape[0], profile=profile).prove(blocks, domain=domain)[0]


def test_mix_equals_alone():
    jobs = [(200, 1, pv.DOMAIN), (1000, 2, pv.DOMAIN), (3000, 3, b"another-domain"), (256, 4, pv.DOMAIN),
            (700, 5, pv.DOMAIN), (1024, 6, b"")]
    inputs = [(_blocks(n, sd), dom) for n, sd, dom in jobs]
    torch.cuda.synchronize()
    pool = pv.ProverPool(workers=3, hbm_budget=8 << 30)
    for n, _, _ in jobs:
        pool.prepare(n)
    tickets = [pool.submit(b, domain=d) for b, d in inputs]
    got = [pool.wait(t) for t in tickets]
    st = pool.stats()
    pool.close()
    assert st["completed"] == len(jobs) and st["peak_reserved"] <= st["budget"]
    for (b, d), (proof, times) in zip(inputs, got):
        print(f"case n={b.shape[0]} domain={d!r}", flush=True)
        assert proof == _alone(b, d)
        assert times[0] >= 0 and times[1] > 0 and abs(times[2] - times[0] - times[1]) < 1e-3


def test_tight_budget_serialises_jobs():
    inputs = [_blocks(n, 10 + n) for n in (900, 1000, 300)]
    torch.cuda.synchronize()
    biggest = pv.Blake3Setup(1000)  # m = 24
    pool = pv.ProverPool(workers=3, hbm_budget=biggest.arena_stats()["capacity"] + (1 << 20))
    got = [pool.wait(t)[0] for t in [pool.submit(b) for b in inputs]]
    st = pool.stats()
    pool.close()
    assert st["peak_running"] == 1 and st["peak_reserved"] <= st["budget"]
    for b, proof in zip(inputs, got):
        assert proof == biggest.prove(b)[0] if b.shape[0] > 512 else proof == _alone(b)


def test_job_over_budget_fails_alone():
    small, big = _blocks(256, 21), _blocks(4096, 22)
    torch.cuda.synchronize()
    cap = pv.Blake3Setup(256).arena_stats()["capacity"]
    pool = pv.ProverPool(workers=2, hbm_budget=cap + (1 << 20))
    tb, ts = pool.submit(big), pool.submit(small)
    with pytest.raises(pv.nat.FlockGpuError):
        pool.wait(tb)
    proof, _ = pool.wait(ts)
    pool.close()
    assert proof == _alone(small)
    assert pv.nat.load().flk_is_disabled() == 0  # a refused job latches nothing


def test_profiles_in_pools():
    b = _blocks(300, 31)
    torch.cuda.synchronize()
    pool = pv.ProverPool(profile="slim", workers=2, hbm_budget=4 << 30)
    proof, _ = pool.wait(pool.submit(b))
    pool.close()
    assert proof == _alone(b, profile="slim")
