"""Seeded synthetic inputs for the benchmark workloads and tests.

One generator, SplitMix64, so every consumer (GPU run, oracle, CPU baseline)
sees the same numbers for a seed:
    state_i = seed + (i + 1) * 0x9E3779B97F4A7C15        (mod 2^64)
    z = mix(state_i); u_i = (z >> 11) * 2^-53           in [0, 1)
Stream order for a pair (x, k) of N complex values: x.re, x.im, k.re, k.im,
i.e. draws [0, N), [N, 2N), [2N, 3N), [3N, 4N).
U(-1, 1) = 2u - 1; N(0, 1) by Box-Muller on consecutive draw pairs.
"""
from __future__ import annotations

import numpy as np

_G = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def uniform01(seed: int, start: int, count: int) -> np.ndarray:
    i = np.arange(start + 1, start + count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * _G
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_pm1(seed: int, start: int, count: int) -> np.ndarray:
    return 2.0 * uniform01(seed, start, count) - 1.0


def normal(seed: int, start: int, count: int) -> np.ndarray:
    m = (count + 1) // 2
    u = uniform01(seed, start, 2 * m)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    th = 2.0 * np.pi * u2
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(th)
    out[1::2] = r * np.sin(th)
    return out[:count]


def complex_uniform(seed: int, shape, slot: int = 0) -> np.ndarray:
    """Complex array with re, im i.i.d. U(-1,1): draws [2 slot N, 2 slot N + 2N)."""
    n = int(np.prod(shape))
    re = uniform_pm1(seed, 2 * slot * n, n)
    im = uniform_pm1(seed, 2 * slot * n + n, n)
    return (re + 1j * im).reshape(shape)


def complex_normal(seed: int, shape, slot: int = 0) -> np.ndarray:
    n = int(np.prod(shape))
    re = normal(seed, 2 * slot * n, n)
    im = normal(seed, 2 * slot * n + n, n)
    return (re + 1j * im).reshape(shape)


def gaussian_stencil_3d(n: int, sigma: float = 3.0) -> np.ndarray:
    """k = g(i1) g(i2) g(i3), g(i) = exp(-d(i)^2 / (2 sigma^2)), d(i) = min(i, n - i); sum k = 1, imag 0."""
    i = np.arange(n)
    d = np.minimum(i, n - i).astype(np.float64)
    g = np.exp(-(d * d) / (2.0 * sigma * sigma))
    g /= g.sum()
    return (g[:, None, None] * g[None, :, None] * g[None, None, :]).astype(np.complex128)


# ---- the five workloads ------------------------------------------------------
def workload_inputs(name: str):
    """(x, k) complex128 arrays of a named workload ("batched_4096" returns lists of pairs)."""
    if name == "1d_65536":  # 1D 2^16, U(-1,1), seed 0
        return complex_uniform(0, (1 << 16,), 0), complex_uniform(0, (1 << 16,), 1)
    if name == "2d_2048x2048":  # 2D 2048^2, U(-1,1), seed 1
        return complex_uniform(1, (2048, 2048), 0), complex_uniform(1, (2048, 2048), 1)
    if name == "3d_256cubed":  # 3D 256^3, x ~ N(0,1), Gaussian stencil sigma = 3, seed 2
        return complex_normal(2, (256, 256, 256), 0), gaussian_stencil_3d(256, 3.0)
    if name == "2d_4096x256":  # 2D 4096 x 256, U(-1,1), seed 3
        return complex_uniform(3, (4096, 256), 0), complex_uniform(3, (4096, 256), 1)
    if name == "batched_4096":  # 4096 pairs of lengths 2^U{8..14}, U(-1,1), seed 4
        return batched_inputs(4, 4096)
    raise ValueError(name)


def batched_lengths(seed: int, count: int, lo: int = 8, hi: int = 14) -> np.ndarray:
    """m_i uniform on {lo..hi} from draws [0, count) of the seed's stream."""
    u = uniform01(seed, 0, count)
    return lo + np.minimum((u * (hi - lo + 1)).astype(np.int64), hi - lo)


def batched_inputs(seed: int, count: int):
    ms = batched_lengths(seed, count)
    xs, ks = [], []
    off = count
    for m in ms:
        n = 1 << int(m)
        xr = uniform_pm1(seed, off, n)
        xi = uniform_pm1(seed, off + n, n)
        kr = uniform_pm1(seed, off + 2 * n, n)
        ki = uniform_pm1(seed, off + 3 * n, n)
        off += 4 * n
        xs.append(xr + 1j * xi)
        ks.append(kr + 1j * ki)
    return xs, ks
