"""Deterministic synthetic scalar volumes (float32, x-fastest raw layout).

Every workload in this package is generated here from a fixed seed, so the
same bytes are produced on any machine with this numpy/scipy build:

* counter-based SplitMix64 stream -> uniform doubles in [0, 1) (53-bit),
* Box-Muller in float64 for normal deviates,
* arithmetic in float64, one final cast to float32.

Arrays are returned with shape (Z, Y, X) so that ``a.ravel()`` is the
reference's vertex order ``x + X*y + X*Y*z`` (gi_regular_grid.h:54-56).
"""
from __future__ import annotations

import hashlib

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started at ``seed``; element i is the (offset+i+1)-th output."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


class Stream:
    """Sequential view over the SplitMix64 stream of one seed."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.pos = 0

    def uniform(self, n: int) -> np.ndarray:
        z = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def uniform_range(self, lo: float, hi: float, n: int) -> np.ndarray:
        return lo + (hi - lo) * self.uniform(n)

    def normal(self, n: int) -> np.ndarray:
        u1 = 1.0 - self.uniform(n)  # (0, 1]
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)


def gaussian_mixture(shape_zyx, seed: int, count: int, sigma_lo: float, sigma_hi: float,
                     amp_lo: float = -1.0, amp_hi: float = 1.0, noise: float = 0.0,
                     sine: float = 0.0, k_lo: float = 0.05, k_hi: float = 0.2) -> np.ndarray:
    """Sum of isotropic Gaussians (+ optional uniform noise and a sin*sin*sin term)."""
    Z, Y, X = shape_zyx
    s = Stream(seed)
    cx = s.uniform_range(0.0, X - 1.0, count)
    cy = s.uniform_range(0.0, Y - 1.0, count)
    cz = s.uniform_range(0.0, Z - 1.0, count)
    sig = s.uniform_range(sigma_lo, sigma_hi, count)
    amp = s.uniform_range(amp_lo, amp_hi, count)
    x = np.arange(X, dtype=np.float64)
    y = np.arange(Y, dtype=np.float64)
    z = np.arange(Z, dtype=np.float64)
    f = np.zeros((Z, Y, X), dtype=np.float64)
    for g in range(count):
        inv = -0.5 / (sig[g] * sig[g])
        ex = np.exp(inv * (x - cx[g]) ** 2)
        ey = np.exp(inv * (y - cy[g]) ** 2)
        ez = np.exp(inv * (z - cz[g]) ** 2)
        f += amp[g] * (ez[:, None, None] * (ey[None, :, None] * ex[None, None, :]))
    if sine != 0.0:
        kx, ky, kz = s.uniform_range(k_lo, k_hi, 3)
        f += sine * (np.sin(kz * z)[:, None, None] * np.sin(ky * y)[None, :, None]
                     * np.sin(kx * x)[None, None, :])
    if noise != 0.0:
        f += s.uniform_range(-noise, noise, Z * Y * X).reshape(Z, Y, X)
    return f.astype(np.float32)


def smoothed_noise(shape_zyx, seed: int, sigma: float) -> np.ndarray:
    """N(0,1) white noise filtered by a separable Gaussian (truncate 4 sigma, mirror edges)."""
    from scipy.ndimage import gaussian_filter

    Z, Y, X = shape_zyx
    s = Stream(seed)
    n = s.normal(Z * Y * X).reshape(Z, Y, X)
    out = gaussian_filter(n, sigma=sigma, mode="mirror", truncate=4.0)
    return out.astype(np.float32)


# Named workloads (sizes and seeds of the benchmark and parity configurations).
def lstar_small(seed: int = 101) -> np.ndarray:
    """64^3: 32 Gaussians (centres in the box, sigma 3-10, amplitude +-1) + U(+-1e-3)."""
    return gaussian_mixture((64, 64, 64), seed, 32, 3.0, 10.0, -1.0, 1.0, noise=1e-3)


def lstar_large(seed: int = 202, n: int = 512) -> np.ndarray:
    """n^3 (512^3 by default): N(0,1) noise smoothed with sigma = 2 voxels."""
    return smoothed_noise((n, n, n), seed, 2.0)


def labels_field(seed: int = 303, n: int = 256) -> np.ndarray:
    """n^3 (256^3 by default): 200 Gaussians (sigma 4-16) + 0.05 sin(kx x) sin(ky y) sin(kz z)."""
    return gaussian_mixture((n, n, n), seed, 200, 4.0, 16.0, -1.0, 1.0, sine=0.05)


def sha256(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def batch_volume(i: int) -> np.ndarray:
    """Volume i of the batched workload (seed 4000 + i): sides uniform in
    32..160 per axis; smoothed noise (sigma 1.5-3) or a Gaussian mixture
    (200 Gaussians per 256^3, sigma 4-16, with the sine term), 50/50."""
    s = Stream(4000 + i)
    u = s.uniform(5)
    shape = tuple(int(32 + min(128, int(129 * float(u[k])))) for k in range(3))  # (z, y, x)
    if u[3] < 0.5:
        return smoothed_noise(shape, 4000 + i, 1.5 + 1.5 * float(u[4]))
    n = max(4, int(round(200 * shape[0] * shape[1] * shape[2] / 256 ** 3)))
    return gaussian_mixture(shape, 4000 + i, n, 4.0, 16.0, -1.0, 1.0, sine=0.05)
