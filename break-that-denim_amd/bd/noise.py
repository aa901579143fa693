"""Deterministic integer hashing and smooth noise for scene compilation (vectorised numpy).

Restates the seeded generators of the reference's asset compiler
(src/shared/noise.ts:4-80): the lowbias32 integer hash, a 3-argument hash chain, uniform
draws from it, 2D gradient noise with a quintic fade, its fractal sum and a 1D yarn noise.
All functions accept numpy arrays (broadcast) or Python scalars.
"""
from __future__ import annotations

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def _u32(x) -> np.ndarray:
    """ToUint32 of integer-valued input (two's complement wrap, like JS `x >>> 0`)."""
    a = np.asarray(x)
    if a.dtype.kind == "f":
        a = a.astype(np.int64)
    return (a.astype(np.int64) & 0xFFFFFFFF).astype(np.uint64)


def hash32(x) -> np.ndarray:
    """lowbias32 (noise.ts:4-12); returns uint64 holding the unsigned 32-bit result."""
    x = _u32(x)
    x ^= x >> np.uint64(16)
    x = (x * np.uint64(0x7FEB352D)) & _M32
    x ^= x >> np.uint64(15)
    x = (x * np.uint64(0x846CA68B)) & _M32
    x ^= x >> np.uint64(16)
    return x


def hash3(a, b, c) -> np.ndarray:
    """hash32(a ^ hash32(b ^ hash32(c))) on 32-bit words (noise.ts:14-16)."""
    return hash32(_u32(a) ^ hash32(_u32(b) ^ hash32(c)))


def rand01(a, b=0, c=0):
    """Uniform in [0, 1) (noise.ts:19-21)."""
    return hash3(a, b, c).astype(np.float64) / 4294967296.0


def _fade(t):
    return t * t * t * (t * (t * 6 - 15) + 10)


def _grad(seed, ix, iy, fx, fy):
    h = hash3(seed, ix, iy).astype(np.float64)
    ang = (h / 4294967296.0) * np.pi * 2
    return np.cos(ang) * fx + np.sin(ang) * fy


def gradient_noise(seed, x, y):
    """2D gradient noise, approximately in [-0.7, 0.7] (noise.ts:38-52)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    ix = np.floor(x)
    iy = np.floor(y)
    fx = x - ix
    fy = y - iy
    u = _fade(fx)
    v = _fade(fy)
    ixi = ix.astype(np.int64)
    iyi = iy.astype(np.int64)
    n00 = _grad(seed, ixi, iyi, fx, fy)
    n10 = _grad(seed, ixi + 1, iyi, fx - 1, fy)
    n01 = _grad(seed, ixi, iyi + 1, fx, fy - 1)
    n11 = _grad(seed, ixi + 1, iyi + 1, fx - 1, fy - 1)
    a = n00 + (n10 - n00) * u
    b = n01 + (n11 - n01) * u
    return a + (b - a) * v


def fbm(seed, x, y, octaves, lacunarity=2.03, gain=0.5):
    """Fractal sum of gradient noise, roughly in [-1, 1] (noise.ts:55-69)."""
    s = 0.0
    amp = 1.0
    norm = 0.0
    fx = np.asarray(x, dtype=np.float64)
    fy = np.asarray(y, dtype=np.float64)
    for o in range(octaves):
        s = s + amp * gradient_noise(seed + o * 1013, fx, fy)
        norm += amp
        amp *= gain
        fx = fx * lacunarity
        fy = fy * lacunarity
    return (s / norm) * 1.6


def noise1(seed, x):
    """1D smooth noise along a yarn (noise.ts:72-74); `seed` may be an integer array."""
    seed = np.asarray(seed, dtype=np.int64)
    return gradient_noise(seed, x, 0.5 + (seed % 7) * 0.13)


def smoothstep(e0, e1, x):
    with np.errstate(divide="ignore", invalid="ignore"):  # e0 == e1 gives NaN/inf like JS
        t = np.clip((np.asarray(x, dtype=np.float64) - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3 - 2 * t)
