"""Seeded synthetic inputs of the benchmark workloads.

SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) drives every random choice so the
same seed gives the same problem on any machine; normals use Box-Muller.
"""
from __future__ import annotations

import math

import numpy as np

MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """uniform in [0, 1) with 53 random bits"""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def randint(self, lo: int, hi: int) -> int:
        """uniform integer in [lo, hi]"""
        return lo + self.next_u64() % (hi - lo + 1)

    def normal(self) -> float:
        u1 = max(self.uniform(), 1e-300)
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def sinusoid_terms(seed: int, nterms: int, dim: int, kmax: int = 8):
    """(kw[nterms, dim], amp[nterms]): integer wavenumbers 1..kmax, N(0,1) amplitudes."""
    rng = SplitMix64(seed)
    kw = np.zeros((nterms, dim), dtype=np.int32)
    amp = np.zeros(nterms)
    for j in range(nterms):
        for a in range(dim):
            kw[j, a] = rng.randint(1, kmax)
        amp[j] = rng.normal()
    return kw, amp


def splitmix_normals(seed: int, count: int) -> np.ndarray:
    """count standard normals from the SplitMix64 stream of `seed` (vectorised:
    output i is the mix of seed + (i+1)*golden), one Box-Muller normal per
    consecutive pair of uniforms."""
    m = 2 * count
    i = np.arange(1, m + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK) + i * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    u1 = np.maximum(u[0::2], 1e-300)
    u2 = u[1::2]
    return (np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2))[:count]


def se_covariance_1d(n: int, length: float) -> np.ndarray:
    """squared-exponential covariance exp(-(x-y)^2 / (2 length^2)) between the n cell
    centres (i + 1/2) / n of the unit interval"""
    x = (np.arange(n) + 0.5) / n
    return np.exp(-((x[:, None] - x[None, :]) ** 2) / (2 * length ** 2))


def gaussian_field_raw(seed: int, cells, length: float = 0.1) -> np.ndarray:
    """an exact sample of the zero-mean Gaussian field with covariance
    K_0 (x) K_1 (x) ... (K_a = se_covariance_1d(cells[a], length)) over the cell centres
    of the unit box, flattened in the reference's cell order (last axis fastest):
    g = (S_0 (x) S_1 (x) ...) xi with S_a the symmetric square root of K_a
    (eigendecomposition) and xi the SplitMix64 normals of `seed`."""
    cells = tuple(int(c) for c in cells)
    g = splitmix_normals(seed, int(np.prod(cells))).reshape(cells)
    for a, n in enumerate(cells):
        lam, V = np.linalg.eigh(se_covariance_1d(n, length))
        S = (V * np.sqrt(np.maximum(lam, 0.0))) @ V.T
        g = np.moveaxis(np.tensordot(S, g, axes=([1], [a])), 0, a)
    return g.reshape(-1)


def gaussian_random_field(seed: int, cells, length: float = 0.1, contrast: float = 1e3):
    """log-coefficient g per cell (reference cell order: last axis fastest): the sample
    gaussian_field_raw(seed, cells, length), centred and rescaled so that
    max g - min g = ln(contrast).  The coefficient a = exp(g) then has
    max a / min a = contrast exactly."""
    g = gaussian_field_raw(seed, cells, length)
    g = g - g.mean()
    span = g.max() - g.min()
    if span > 0:
        g = g * (np.log(contrast) / span)
    return g


# the five workloads: name -> (dim, degree, cells, start, end, seed, nterms)
CONFIGS = {
    "c1_2d_q1_64": dict(dim=2, degree=1, cells=(64, 64), seed=1, nterms=8),
    "c2_3d_q2_96": dict(dim=3, degree=2, cells=(96, 96, 96), seed=2, nterms=16),
    "c4_2d_q1_1024": dict(dim=2, degree=1, cells=(1024, 1024), seed=4, nterms=8),
}
