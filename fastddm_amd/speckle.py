"""Seeded synthetic Brownian-speckle image stacks for tests and benchmarks.

Each wavevector q of the (H, W) grid (2 pi fftfreq per axis) carries a complex
Gaussian amplitude with envelope exp(-|q|^2 sigma^2 / 2).  It evolves as an
AR(1) process with rho(q) = exp(-D |q|^2) and innovation variance 1 - rho^2,
so the field decorrelates like Brownian motion with diffusivity D
(px^2/frame).  Every frame is real(ifft2(amplitude)).  The stack is then
mapped affinely to the target mean / std (statistics taken over the whole
stack), optionally Poisson-sampled, clipped to the dtype range and cast.
"""

import os
from typing import Optional

import numpy as np
from scipy import fft as sfft


def _workers() -> int:
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def ar1_speckle(
    length: int,
    height: int,
    width: int,
    *,
    sigma: float = 2.0,
    diffusivity: float = 0.5,
    mean: float = 2000.0,
    std: float = 300.0,
    dtype=np.uint16,
    seed: int = 1,
    poisson: bool = True,
    clip: Optional[tuple] = None,
) -> np.ndarray:
    """Return an (length, height, width) stack of `dtype` (see module doc)."""
    rng = np.random.default_rng(seed)
    ky = 2.0 * np.pi * np.fft.fftfreq(height)
    kx = 2.0 * np.pi * np.fft.fftfreq(width)
    q2 = ky[:, None] ** 2 + kx[None, :] ** 2
    env = np.exp(-0.5 * q2 * sigma**2)
    rho = np.exp(-diffusivity * q2)
    innov = np.sqrt(np.maximum(0.0, 1.0 - rho**2)) * env

    def noise():
        return (rng.standard_normal((height, width)) + 1j * rng.standard_normal((height, width))) / np.sqrt(2.0)

    amp = noise() * env
    frames = np.empty((length, height, width), dtype=np.float32)
    workers = _workers()
    for t in range(length):
        frames[t] = sfft.ifft2(amp, workers=workers).real
        amp = rho * amp + innov * noise()
    m = float(frames.mean(dtype=np.float64))
    s = float(frames.std(dtype=np.float64))
    scale = std / s if s > 0 else 0.0
    frames -= m
    frames *= scale
    frames += mean
    if poisson:
        np.clip(frames, 0.0, None, out=frames)
        frames = rng.poisson(frames).astype(np.float64)
    if np.issubdtype(np.dtype(dtype), np.integer):
        info = np.iinfo(dtype)
        lo, hi = (clip if clip is not None else (info.min, info.max))
        np.clip(frames, lo, hi, out=frames)
        return np.rint(frames).astype(dtype)
    if clip is not None:
        np.clip(frames, clip[0], clip[1], out=frames)
    return frames.astype(dtype)


# The five workloads the benchmarks and tests draw from (sizes, dynamics and
# intensity statistics of each stack).
CONFIGS = {
    "cpu_128": dict(length=64, height=128, width=128, sigma=2.0, diffusivity=0.5,
                    mean=2000.0, std=300.0, dtype=np.uint16, seed=1),
    "precision_512": dict(length=2000, height=512, width=512, sigma=1.5, diffusivity=0.2,
                          mean=1000.0, std=150.0, dtype=np.uint16, seed=2),
    "diff_256": dict(length=300, height=256, width=256, sigma=2.0, diffusivity=1.0,
                     mean=120.0, std=25.0, dtype=np.uint8, seed=3, clip=(0, 255)),
}
