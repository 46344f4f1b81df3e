"""Seeded synthetic inputs: SplitMix64 → U[0, 1) in FP64, identical for every consumer.

Element n (0-based, x fastest over the field interior) of stream `seed` is the SplitMix64
output for state seed + (n + 1)·0x9E3779B97F4A7C15, mapped to [0, 1) with its top 53 bits.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, n):
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed, shape, lo=0.0, hi=1.0):
    """U[lo, hi) samples of the given (Nx, Ny, Nz) shape, x fastest in the stream."""
    n = int(np.prod(shape))
    u = (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return (lo + (hi - lo) * u).reshape(shape, order="F")
