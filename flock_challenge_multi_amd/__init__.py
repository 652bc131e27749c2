"""MI355X-native (gfx950) Flock prover hot path.

The product is the C-ABI device library `libflock_gpu.so` (include/flock_gpu.h,
sources in csrc/).  `hotpath` mirrors the reference's Rust entry points over it
for Python callers; `_native` is the raw ctypes binding.
"""
from ._native import FlockGpuError, load  # noqa: F401

__all__ = ["load", "FlockGpuError"]
