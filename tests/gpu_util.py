"""Helpers for GPU parity tests: numpy F128 arrays <-> device tensors."""
import numpy as np
import torch

from oracle_ffi import F128_DT


def to_dev(a):
    """numpy F128_DT array -> contiguous CUDA int64 tensor (n, 2)."""
    v = np.ascontiguousarray(a).view(np.int64).reshape(-1, 2)
    return torch.from_numpy(v.copy()).cuda()


def from_dev(t):
    """CUDA int64 (n, 2) tensor -> numpy F128_DT array."""
    h = t.detach().cpu().numpy().astype(np.int64, copy=False)
    return np.ascontiguousarray(h).view(F128_DT).reshape(-1)


def bytes_from_dev(t):
    return t.detach().cpu().numpy().astype(np.uint8).reshape(-1, 32)


def same(a, b):
    return a.shape == b.shape and bool((a["lo"] == b["lo"]).all() and (a["hi"] == b["hi"]).all())
