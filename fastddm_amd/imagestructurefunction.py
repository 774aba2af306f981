"""Result container of ``ddm()``.

Same data contract as the reference ``ImageStructureFunction``
(src/python/imagestructurefunction.py:71-406): ``_data`` has shape
``(NL + 2, ny, nx // 2 + 1)``; the first NL slices are D(q, dt) per lag, slice
``-2`` is the average power spectrum and slice ``-1`` the variance over time,
all fftshifted along y. ``kx`` is the non-negative half axis, ``ky`` the full
shifted axis, ``tau`` the lags. ``full_*`` rebuilds the symmetric full plane
from the half plane (reference ``_reconstruct_full_spectrum``, ``:887-944``).

``save`` writes the reference's ``.sf.ddm`` binary format (``fastddm_amd.sfio``).
"""

from dataclasses import dataclass
from typing import Tuple

import numpy as np


@dataclass
class ImageStructureFunction:
    """Image structure function container.

    Parameters
    ----------
    _data : numpy.ndarray
        Packed data ``(NL + 2, ny, nx // 2 + 1)``: D slices, power spectrum, variance.
    _kx, _ky : numpy.ndarray
        Wave-vector axes.
    _width, _height : int
        Width and height of the full (symmetric) plane, i.e. ``nx`` and ``ny``.
    _tau : numpy.ndarray
        Time delays.
    _pixel_size, _delta_t : float
        Effective pixel size and frame interval (default 1).
    """

    _data: np.ndarray
    _kx: np.ndarray
    _ky: np.ndarray
    _width: int
    _height: int
    _tau: np.ndarray
    _pixel_size: float = 1.0
    _delta_t: float = 1.0

    @property
    def data(self) -> np.ndarray:
        """D(q, dt) in ``(dt, ky, kx)`` order."""
        return self._data[:-2]

    @property
    def kx(self) -> np.ndarray:
        return self._kx

    @property
    def ky(self) -> np.ndarray:
        return self._ky

    @property
    def width(self) -> int:
        return self._width

    @property
    def height(self) -> int:
        return self._height

    @property
    def tau(self) -> np.ndarray:
        return self._tau

    @property
    def power_spec(self) -> np.ndarray:
        """Average 2D power spectrum of the input images."""
        return self._data[-2]

    @property
    def var(self) -> np.ndarray:
        """Variance over time of the Fourier transformed images."""
        return self._data[-1]

    @property
    def shape(self) -> Tuple[int, int, int]:
        return self.data.shape

    @property
    def pixel_size(self) -> float:
        return self._pixel_size

    @pixel_size.setter
    def pixel_size(self, pixel_size: float) -> None:
        """Set the pixel size; rescales ``kx`` and ``ky`` accordingly."""
        self._kx *= self._pixel_size / pixel_size
        self._ky *= self._pixel_size / pixel_size
        self._pixel_size = pixel_size

    @property
    def delta_t(self) -> float:
        return self._delta_t

    @delta_t.setter
    def delta_t(self, delta_t: float) -> None:
        """Set the frame interval; rescales ``tau`` accordingly."""
        self._tau *= delta_t / self._delta_t
        self._delta_t = delta_t

    def __len__(self) -> int:
        return len(self.data)

    def set_frame_rate(self, frame_rate: float) -> None:
        self.delta_t = 1 / frame_rate

    def save(self, fname: str = "analysis_blob") -> str:
        """Write to ``fname`` (``.sf.ddm`` appended when missing) in the
        reference's binary format (imagestructurefunction.py:279-293)."""
        from .sfio import save

        return save(self, fname)

    def full_shape(self) -> Tuple[int, int, int]:
        dim_t, dim_y = self.shape[:-1]
        return dim_t, dim_y, self._width

    def full_slice(self, idx: int) -> np.ndarray:
        """Full symmetric plane of D at lag index ``idx``."""
        if 0 <= idx < len(self.data):
            return _reconstruct_full_spectrum(self._data[idx], (self.height, self.width))
        raise IndexError(
            f"Index out of range. Choose an index between 0 and {len(self.data)}."
        )

    def full_power_spec(self) -> np.ndarray:
        return _reconstruct_full_spectrum(self._data[-2], (self.height, self.width))

    def full_var(self) -> np.ndarray:
        return _reconstruct_full_spectrum(self._data[-1], (self.height, self.width))

    def full_kx(self) -> np.ndarray:
        """Full, shifted ``kx`` axis."""
        full = np.zeros_like(self.kx, shape=(self.width,))
        half = len(self.kx)
        full[:half] = self.kx
        if self.width % 2 == 0:
            full[half:] = -np.flip(self.kx[1:-1])
        else:
            full[half:] = -np.flip(self.kx[1:])
        return np.fft.fftshift(full)

    def full_ky(self) -> np.ndarray:
        return self.ky


def _reconstruct_full_spectrum(halfplane: np.ndarray, shape: Tuple[int, int]) -> np.ndarray:
    """Full symmetric plane (shifted along both axes) from a y-shifted half plane.

    The half plane holds columns kx = 0 .. nx//2 with rows already fftshifted
    along y. The missing columns are the point reflection of columns
    1 .. (nx - 1)//2 through the centre of the shifted grid; for even ny the
    first shifted row (ky = -ny/2) is its own mirror and is reflected in place.
    """
    height, width = shape
    dim_x = halfplane.shape[1]
    full = np.zeros(shape, dtype=halfplane.dtype)
    full[:, :dim_x] = halfplane
    x_end = dim_x - 1 if width % 2 == 0 else dim_x
    y_start = 1 if height % 2 == 0 else 0
    full[y_start:, dim_x:] = halfplane[y_start:, 1:x_end][::-1, ::-1]
    if height % 2 == 0:
        full[0, dim_x:] = halfplane[0, 1:x_end][::-1]
    return np.fft.fftshift(full, axes=-1)
