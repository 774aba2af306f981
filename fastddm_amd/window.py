"""Window functions applied to the images before the spatial FFT.

Same definitions as the reference (src/python/window.py:54-161): periodic
cosine-sum windows w(x) = sum_j (-1)^j a_j cos(2 pi j x / N), 0 <= x < N,
evaluated separately along x and y; the 2D window is the outer product
W(y, x) = w_y(y) w_x(x), cast to ``DTYPE``.
"""

from typing import Sequence, Tuple

import numpy as np

from ._config import DTYPE

# Coefficients: exact Blackman (reference window.py:93) and the four-term
# Blackman-Harris set (reference window.py:149).
_BLACKMAN = (7938.0 / 18608.0, 9240.0 / 18608.0, 1430.0 / 18608.0)
_BLACKMAN_HARRIS = (0.3635819, 0.4891775, 0.1365995, 0.0106411)


def _cosine_sum_1d(n: int, coeffs: Sequence[float]) -> np.ndarray:
    x = np.arange(n, dtype=np.float64) / n
    w = np.zeros(n, dtype=np.float64)
    sign = 1.0
    for j, a in enumerate(coeffs):
        w += sign * a * np.cos(2.0 * np.pi * j * x)
        sign = -sign
    return w


def _cosine_sum_2d(shape: Tuple[int, ...], coeffs: Sequence[float]) -> np.ndarray:
    *_, ydim, xdim = shape
    wy = _cosine_sum_1d(ydim, coeffs)
    wx = _cosine_sum_1d(xdim, coeffs)
    return np.multiply.outer(wy, wx).astype(DTYPE)


def blackman(shape: Tuple[int, ...]) -> np.ndarray:
    """Periodic three-term exact Blackman window of the last two dims of ``shape``."""
    return _cosine_sum_2d(shape, _BLACKMAN)


def blackman_harris(shape: Tuple[int, ...]) -> np.ndarray:
    """Periodic four-term Blackman-Harris window of the last two dims of ``shape``."""
    return _cosine_sum_2d(shape, _BLACKMAN_HARRIS)
