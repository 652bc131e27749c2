"""The harness input generator (flock_challenge_multi_amd/inputs.py) against a
scalar restatement of benchmark-tools/common/src/lib.rs:20-34 written here."""
from flock_challenge_multi_amd.inputs import generate_compressions


def _scalar(log2_size, seed, n):
    mask = (1 << 64) - 1
    s = (seed ^ (((log2_size << 29) | (log2_size >> 35)) & mask)) & mask
    out = []
    for _ in range(n):
        rec = []
        for _ in range(25):
            s = (s + 0x9E3779B97F4A7C15) & mask
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
            rec.append((z ^ (z >> 31)) & 0xFFFFFFFF)
        out.append(rec[:24] + [rec[24], 0, 64, 11])
    return out


def test_generator_matches_scalar():
    for log2, seed in ((3, 0), (5, 0xC0FFEE), (12, 2**64 - 1)):
        got = generate_compressions(log2, seed)
        assert got.shape == (1 << log2, 28)
        want = _scalar(log2, seed, min(1 << log2, 40))
        assert [list(map(int, r)) for r in got[: len(want)]] == want
