"""Benchmark inputs of the reference harness.

generate_compressions (benchmark-tools/common/src/lib.rs:20-34): 2^log2_size
BLAKE3 compressions from a SplitMix64 stream seeded with
seed ^ rotl64(log2_size, 29); each draws 8 chaining-value words, 16 message
words and a counter (one 32-bit draw each, the low half of the SplitMix64
output), with block_len = 64 and flags = 11.  Returned as a (n, 28) uint32
array in the 112-byte record layout of flk_compression (cv, m, counter lo,
counter hi, block_len, flags).
"""
import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _rotl64(x, r):
    x &= (1 << 64) - 1
    return ((x << r) | (x >> (64 - r))) & ((1 << 64) - 1)


def generate_compressions(log2_size, seed):
    n = 1 << log2_size
    state0 = np.uint64((seed ^ _rotl64(log2_size, 29)) & ((1 << 64) - 1))
    k = np.arange(1, 25 * n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = state0 + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    draws = (z & np.uint64(0xFFFFFFFF)).astype(np.uint32).reshape(n, 25)
    out = np.zeros((n, 28), np.uint32)
    out[:, :24] = draws[:, :24]
    out[:, 24] = draws[:, 24]  # counter = u64::from(next_u32): high word 0
    out[:, 26] = 64
    out[:, 27] = 11
    return out
