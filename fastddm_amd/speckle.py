"""Seeded synthetic Brownian-speckle image stacks for tests and benchmarks.

Each wavevector q of the (H, W) grid (2 pi fftfreq per axis) carries a complex
Gaussian amplitude with envelope exp(-|q|^2 sigma^2 / 2).  It evolves as an
AR(1) process with rho(q) = exp(-D |q|^2) and innovation variance 1 - rho^2,
so the field decorrelates like Brownian motion with diffusivity D
(px^2/frame).  Every frame is real(ifft2(amplitude)).  The stack is then
mapped affinely to the target mean / std (statistics taken over the whole
stack), optionally Poisson-sampled, clipped to the dtype range and cast.

Randomness: ``SeedSequence(seed)`` spawns one child stream for the initial
amplitude and one per frame (innovation, then shot noise), so the stack is a
pure function of the seed whatever the number of worker threads.
"""

import os
from concurrent.futures import ThreadPoolExecutor
from typing import Optional

import numpy as np
from scipy import fft as sfft


def _workers() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


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
    kids = np.random.SeedSequence(seed).spawn(length + 1)
    ky = 2.0 * np.pi * np.fft.fftfreq(height)
    kx = 2.0 * np.pi * np.fft.fftfreq(width)
    q2 = ky[:, None] ** 2 + kx[None, :] ** 2
    env = np.exp(-0.5 * q2 * sigma**2).astype(np.float32)
    rho = np.exp(-diffusivity * q2).astype(np.float32)
    innov = (np.sqrt(np.maximum(0.0, 1.0 - np.exp(-diffusivity * q2) ** 2)) * env).astype(np.float32)
    shape = (height, width)
    scale = np.float32(1.0 / np.sqrt(2.0))

    def cnoise(rng):
        re = rng.standard_normal(shape, dtype=np.float32)
        im = rng.standard_normal(shape, dtype=np.float32)
        return (re + 1j * im).astype(np.complex64) * scale

    workers = _workers()
    frames = np.empty((length, height, width), dtype=np.float32)
    amp = cnoise(np.random.default_rng(kids[0])) * env
    chunk = 64
    with ThreadPoolExecutor(workers) as ex:
        for c0 in range(0, length, chunk):
            c1 = min(length, c0 + chunk)
            eps = list(ex.map(lambda k: cnoise(np.random.default_rng(k)), kids[c0 + 1 : c1 + 1]))
            amps = np.empty((c1 - c0, height, width), dtype=np.complex64)
            for i in range(c1 - c0):
                amps[i] = amp
                amp = rho * amp + innov * eps[i]
            frames[c0:c1] = sfft.ifft2(amps, workers=workers).real
        m = float(frames.mean(dtype=np.float64))
        s = float(frames.std(dtype=np.float64))
        a = np.float32(std / s if s > 0 else 0.0)
        b = np.float32(mean - m * (std / s if s > 0 else 0.0))
        is_int = np.issubdtype(np.dtype(dtype), np.integer)
        if is_int:
            info = np.iinfo(dtype)
            lo, hi = clip if clip is not None else (info.min, info.max)
        else:
            lo, hi = clip if clip is not None else (-np.inf, np.inf)
        out = np.empty((length, height, width), dtype=dtype)

        def finish(t):
            x = frames[t] * a + b
            if poisson:
                rng = np.random.default_rng(kids[t + 1].spawn(1)[0])
                x = rng.poisson(np.clip(x, 0.0, None).astype(np.float64)).astype(np.float64)
            x = np.clip(x, lo, hi)
            out[t] = np.rint(x) if is_int else x

        list(ex.map(finish, range(length)))
    return out


# Workloads the benchmarks and tests draw from (sizes, dynamics and intensity
# statistics of each stack).
CONFIGS = {
    "cpu_128": dict(length=64, height=128, width=128, sigma=2.0, diffusivity=0.5,
                    mean=2000.0, std=300.0, dtype=np.uint16, seed=1),
    "precision_512": dict(length=2000, height=512, width=512, sigma=1.5, diffusivity=0.2,
                          mean=1000.0, std=150.0, dtype=np.uint16, seed=2),
    "diff_256": dict(length=300, height=256, width=256, sigma=2.0, diffusivity=1.0,
                     mean=120.0, std=25.0, dtype=np.uint8, seed=3, clip=(0, 255)),
}
