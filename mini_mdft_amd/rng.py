"""Counter-based random numbers shared by the host, the HIP kernels and the tests.

u_i = (splitmix64(splitmix64(seed) + i) >> 11) * 2**-53 is a uniform double in
[0, 1) that depends only on (seed, i), so numpy, C and HIP code produce the same
stream without sharing state.  The same formula is implemented by
``mdft_fill_uniform`` in the HIP library.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """The splitmix64 finaliser applied element-wise to a uint64 array."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + _GOLDEN).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(30))) * _M1).astype(np.uint64)
        z = ((z ^ (z >> np.uint64(27))) * _M2).astype(np.uint64)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n doubles in [0, 1): element i is a pure function of (seed, offset + i)."""
    key = splitmix64(np.array([seed & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
    with np.errstate(over="ignore"):
        idx = np.arange(offset, offset + n, dtype=np.uint64) + key
    z = splitmix64(idx)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def xi_field(seed: int, shape, amp: float = 0.1) -> np.ndarray:
    """xi = 1 + amp * U(-1, 1), the synthetic density field of the benchmarks,
    written as lo + (hi - lo) u with lo = 1 - amp, hi = 1 + amp (the formula of
    mdft_fill_uniform, so the GPU fill gives the same doubles)."""
    n = int(np.prod(shape))
    lo, hi = 1.0 - amp, 1.0 + amp
    return (lo + (hi - lo) * uniform(seed, n)).reshape(shape)
