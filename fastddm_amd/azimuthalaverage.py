"""Azimuthal average of the image structure function, reduced on the device.

Follows ``azimuthal_average`` / ``azimuthal_average_array`` of the reference
(src/python/azimuthalaverage.py:277-566): bins over |k| (``int`` bins =
``np.linspace(min, max, bins)`` edges with ``np.searchsorted`` digitising, or
explicit monotone edges), optional mask and weights, the half-plane counts
(first column once, last column once if the full width is even, every other
column twice), weighted mean per bin and the unbiased weighted standard
deviation.  Empty bins give NaN and take the bin's edge midpoint as k.

The per-pixel bin indices and weights are prepared here (they depend only on
the q grid); the (rows, ny, nxh) data are reduced by the HIP kernels of
``fddm_azimuthal_average`` where they lie -- a result still in device memory
(``azimuthal_average_device``) never crosses the host link.
"""

import ctypes
import hashlib
import threading
import warnings
from collections import OrderedDict
from dataclasses import dataclass
from typing import Iterable, Optional, Tuple, Union

import numpy as np

from ._config import DTYPE
from .imagestructurefunction import ImageStructureFunction


@dataclass
class AzimuthalAverage:
    """Azimuthal average container (reference ``AzimuthalAverage``,
    azimuthalaverage.py:81-196): ``_data``/``_err`` of shape (bins, NL+2)
    with the power spectrum and variance in the last two columns."""

    _data: np.ndarray
    _err: Optional[np.ndarray]
    k: np.ndarray
    tau: np.ndarray
    bin_edges: np.ndarray

    @property
    def data(self) -> np.ndarray:
        return self._data[:, :-2]

    @property
    def err(self) -> Optional[np.ndarray]:
        return None if self._err is None else self._err[:, :-2]

    @property
    def power_spec(self) -> np.ndarray:
        return self._data[:, -2]

    @property
    def var(self) -> np.ndarray:
        return self._data[:, -1]

    @property
    def power_spec_err(self) -> Optional[np.ndarray]:
        return None if self._err is None else self._err[:, -2]

    @property
    def var_err(self) -> Optional[np.ndarray]:
        return None if self._err is None else self._err[:, -1]

    @property
    def shape(self) -> Tuple[int, int]:
        return self.data.shape


def bin_maps(dist, bins=10, range=None, mask=None, weights=None, counts=None, eval_err=True):
    """Per-pixel bin index and weights, bin edges and the k of every bin.

    Returns ``(bin_index, w, w2, bin_edges, k)``: ``bin_index`` (int32, the
    ``searchsorted`` bin clamped into range), ``w = counts * weights`` and
    ``w2 = counts * weights**2`` with zeros for excluded pixels (masked, below
    the range, above the last edge, zero weight), as the reference's
    ``azimuthal_average_array`` builds them (azimuthalaverage.py:348-440).
    """
    dist = np.asarray(dist)
    if mask is not None:
        if mask.shape != dist.shape:
            raise ValueError("mask shape and dist shape must be the same.")
        _mask = mask.copy()
        if mask.dtype != bool:
            _mask = mask.astype(bool)
            warnings.warn("mask is not of boolean type, it is cast to bool.")
    else:
        _mask = np.full(dist.shape, True)
    if weights is not None:
        if weights.shape != dist.shape:
            raise ValueError("weights shape and dist shape must be the same.")
        if np.any(weights < 0):
            raise ValueError("weights must be non-negative.")
    if counts is not None:
        if counts.shape != dist.shape:
            raise ValueError("counts shape and dist shape must be the same.")
    else:
        counts = np.ones(dist.shape, dtype=DTYPE)
    if range is None:
        x_min, x_max = np.min(dist), np.max(dist)
    else:
        x_min, x_max = range
        if x_min > x_max:
            x_min, x_max = x_max, x_min
    if isinstance(bins, (int, np.integer)):
        bin_edges = np.linspace(x_min, x_max, bins, dtype=DTYPE)
        n_bins = int(bins)
    elif isinstance(bins, Iterable):
        bins = np.asarray(bins)
        if not np.all(bins[1:] >= bins[:-1]):
            raise ValueError("bins must be monotonically increasing.")
        bin_edges = bins.astype(DTYPE)
        n_bins = len(bins)
        x_min = bins[0]
    else:
        raise ValueError("bins must be an int or an iterable.")

    idx = np.searchsorted(bin_edges, dist)
    _mask[(idx == n_bins) | (dist < x_min)] = False
    w = counts.astype(np.float64)
    w2 = counts.astype(np.float64)
    if weights is not None:
        w = w * weights
        w2 = w2 * weights ** 2
    _mask[np.logical_not(w)] = False
    w = np.where(_mask, w, 0.0)
    w2 = np.where(_mask, w2, 0.0)

    # k per bin: weighted mean distance, or the edge midpoint for empty bins
    flat = np.minimum(idx, n_bins - 1).ravel()
    sw = np.bincount(flat, weights=w.ravel(), minlength=n_bins)
    sdw = np.bincount(flat, weights=(w * dist).ravel(), minlength=n_bins)
    with np.errstate(divide="ignore", invalid="ignore"):
        k = np.where(sw > 0, sdw / np.where(sw > 0, sw, 1.0), 0.0)
    for b in np.flatnonzero(sw == 0):
        k[b] = 0.5 * (bin_edges[b - 1] + bin_edges[b]) if b > 0 else bin_edges[0]
    return (np.minimum(idx, n_bins - 1).astype(np.int32), w, w2, bin_edges, k.astype(DTYPE))


def _half_plane_counts(shape, width):
    """Counts of the half-plane columns (azimuthalaverage.py:540-546)."""
    counts = np.full(shape, 2)
    counts[:, 0] = 1
    counts[:, -1] = 1 if width % 2 == 0 else 2
    return counts


def _k_modulus(kx, ky):
    X, Y = np.meshgrid(kx, ky)
    return np.sqrt(X ** 2 + Y ** 2)


def _reduce(ptr, f32, rows, ny, nxh, bin_index, w, w2, nbins, eval_err):
    from . import _core_hip as core

    avg = np.empty((nbins, rows), dtype=np.float64)
    err = np.empty((nbins, rows), dtype=np.float64) if eval_err else None
    bi = np.ascontiguousarray(bin_index, dtype=np.int32)
    wc = np.ascontiguousarray(w, dtype=np.float64)
    w2c = np.ascontiguousarray(w2, dtype=np.float64)
    core._call("fddm_azimuthal_average", ptr, 1 if f32 else 0, rows, ny, nxh, bi.ctypes.data, wc.ctypes.data,
               w2c.ctypes.data, nbins, 1 if eval_err else 0, avg.ctypes.data,
               None if err is None else err.ctypes.data, None)
    return avg, err


_MAPS_CACHE: "OrderedDict[bytes, tuple]" = OrderedDict()
_MAPS_CACHE_SIZE = 8
_MAPS_LOCK = threading.Lock()


def _maps_key(kx, ky, width, bins, range, mask, weights, eval_err) -> Optional[bytes]:
    """Digest of everything the bin maps depend on; None (not cached) for a
    non-boolean mask, whose cast warns on every call."""
    if mask is not None and np.asarray(mask).dtype != bool:
        return None
    h = hashlib.blake2b(digest_size=20)
    for a in (kx, ky, mask, weights, None if isinstance(bins, (int, np.integer)) else bins):
        if a is None:
            h.update(b"-")
        else:
            a = np.ascontiguousarray(a)
            h.update(f"{a.dtype.str}{a.shape}".encode())
            h.update(a.tobytes())
    h.update(repr((int(width), bins if isinstance(bins, (int, np.integer)) else "edges",
                   None if range is None else tuple(float(r) for r in range), bool(eval_err))).encode())
    return h.digest()


def _device_maps(kx, ky, width, ny, nxh, bins, range, mask, weights, eval_err):
    """Bin maps of the half-plane (ny, nxh) grid, kept for the last few
    geometries: repeated averages of results of one size reuse them."""
    key = _maps_key(kx, ky, width, bins, range, mask, weights, eval_err)
    if key is not None:
        with _MAPS_LOCK:
            if key in _MAPS_CACHE:
                _MAPS_CACHE.move_to_end(key)
                return _MAPS_CACHE[key]
    counts = _half_plane_counts((ny, nxh), width)
    maps = bin_maps(_k_modulus(kx, ky), bins, range, mask, weights, counts, eval_err)
    if np.any(np.diff(maps[0], axis=1) < 0):
        raise ValueError("bin indices must not decrease along kx")
    for a in maps:
        a.setflags(write=False)
    if key is not None:
        with _MAPS_LOCK:
            _MAPS_CACHE[key] = maps
            while len(_MAPS_CACHE) > _MAPS_CACHE_SIZE:
                _MAPS_CACHE.popitem(last=False)
    return maps


def azimuthal_average_device(d_data, shape, kx, ky, tau, width, data_dtype=np.float64, bins=10, range=None,
                             mask=None, weights=None, eval_err=True) -> AzimuthalAverage:
    """Azimuthal average of a (NL+2, ny, nxh) result held in device memory
    (a ``DeviceBuffer`` or a raw device pointer), e.g. a plan's output or a
    served request's result."""
    rows, ny, nxh = shape
    bi, w, w2, edges, k = _device_maps(kx, ky, width, ny, nxh, bins, range, mask, weights, eval_err)
    ptr = d_data.ptr if hasattr(d_data, "ptr") else ctypes.c_void_p(d_data)
    avg, err = _reduce(ptr, np.dtype(data_dtype) == np.float32, rows, ny, nxh, bi, w, w2, len(edges), eval_err)
    return AzimuthalAverage(avg.astype(DTYPE), None if err is None else err.astype(DTYPE), k.copy(), np.asarray(tau),
                            edges.copy())


def azimuthal_average(img_str_func: ImageStructureFunction, bins: Optional[Union[int, Iterable[float]]] = 10,
                      range: Optional[Tuple[float, float]] = None, mask: Optional[np.ndarray] = None,
                      weights: Optional[np.ndarray] = None, eval_err: Optional[bool] = True) -> AzimuthalAverage:
    """Azimuthal average of an ``ImageStructureFunction`` (reference
    ``azimuthal_average``, azimuthalaverage.py:481-566), reduced on the
    current device."""
    from . import _core_hip as core

    data = np.ascontiguousarray(img_str_func._data)
    d = core.DeviceBuffer.from_array(data)
    try:
        return azimuthal_average_device(d, data.shape, img_str_func.kx, img_str_func.ky, img_str_func.tau,
                                        img_str_func.width, data.dtype, bins, range, mask, weights, eval_err)
    finally:
        d.free()
