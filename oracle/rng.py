"""Uniform / gaussian streams for the oracle (test infrastructure only).

The reference draws from SPRNG 2.0 (``uniformHao() = sprng()``,
common/randomHao/source/random_hao.cpp:73-76), which is not vendored and not
available here.  Parity therefore treats the uniforms as an *input*: both the
oracle and the GPU path consume the same uniform stream in the reference's
consumption order (one per HS site per alive walker, two per gaussian, one per
comb slot).

The stream is counter based so that the device can evaluate element ``k``
without a sequential state: u(k) = splitmix64 finaliser of
``seed_mix + (k + 1) * 0x9E3779B97F4A7C15`` reduced to 53 bits.  The same
function is implemented in the HIP library (``afqmc_uniform_dev`` in
afqmclab_amd/csrc/cp_kernels.hpp); the two are checked against each other on the GPU.
"""
import numpy as np

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def seed_mix(seed: int) -> int:
    """Map a user seed to the 64-bit stream key (splitmix64 of the seed)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) * _GOLD + _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return int(z)


def uniforms(key: int, start: int, n: int) -> np.ndarray:
    """Return u(start), ..., u(start+n-1) in [0, 1) as float64."""
    k = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(key) + (k + np.uint64(1)) * _GOLD
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class UniformStream:
    """Sequential view of the counter-based stream (the oracle's uniformHao)."""

    def __init__(self, seed: int = 985456376, counter: int = 0):
        self.key = seed_mix(seed)
        self.counter = counter

    def take(self, n: int) -> np.ndarray:
        u = uniforms(self.key, self.counter, n)
        self.counter += n
        return u

    def uniform(self) -> float:
        return float(self.take(1)[0])


class ListStream:
    """A stream that replays given uniforms (host-provided uniforms)."""

    def __init__(self, values):
        self.values = np.asarray(values, dtype=np.float64)
        self.counter = 0

    def take(self, n: int) -> np.ndarray:
        u = self.values[self.counter:self.counter + n]
        if u.size != n:
            raise ValueError("uniform list exhausted")
        self.counter += n
        return u

    def uniform(self) -> float:
        return float(self.take(1)[0])


PI_REF = 3.14159265358979324  # random_hao.cpp:80


def gaussians_from_uniforms(u: np.ndarray) -> np.ndarray:
    """Box-Muller as gaussianHao: pairs (u1, u2) -> sqrt(-2 ln u1) cos(2 pi u2).

    u1 == 0 is replaced by 1e-300 (random_hao.cpp:78-85).  ``u`` holds
    2*n values in consumption order u1_0, u2_0, u1_1, u2_1, ...
    """
    u = np.asarray(u, dtype=np.float64).reshape(-1, 2)
    u1 = np.where(u[:, 0] == 0.0, 1e-300, u[:, 0])
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * PI_REF * u[:, 1])
