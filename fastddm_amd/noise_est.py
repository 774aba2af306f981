"""Camera-noise (B) estimates from an image structure function or its
azimuthal average, as ``estimate_camera_noise`` of the reference
(src/python/noise_est.py:75-841).

Modes: ``"zero"``; ``"min"`` (minimum at the smallest lag); ``"high_q"``
(mean of D over all lags for |k| in [k_min, k_max]); ``"power_spec"`` /
``"var"`` (twice the mean PS / VAR for |k| in [k_min, k_max]); ``"polyfit"``
(constant term of a quadratic fit of the first ``num_points`` lags).
Defaults, warnings and the returned uncertainties follow the reference,
including two of its conventions kept on purpose for identical results: with
a mask, ``"min"`` maps the position inside the masked selection back through
the unmasked (ky, kx) grid (noise_est.py:437-441), and the azimuthal-average
uncertainties divide by the number of selected values including NaN ones
(``np.size`` of the boolean array, :200-206).

The per-pixel quadratic fit of an ``ImageStructureFunction`` runs as one
batched least-squares solve over all (ky, kx) (``np.polyfit`` with 2-D
``y``): the same estimates as the reference's per-pixel loop.
"""

import warnings
from typing import Optional, Tuple, Union

import numpy as np

from ._config import DTYPE
from .azimuthalaverage import AzimuthalAverage
from .imagestructurefunction import ImageStructureFunction


def _as_bool_mask(mask, shape):
    if mask is None:
        return np.full(shape, True)
    if mask.dtype != bool:
        warnings.warn("Given mask not of boolean type. Casting to bool.")
        return mask.astype(bool)
    return mask


def _k_range(k_min, k_max, default):
    k_min = default if k_min is None else k_min
    k_max = default if k_max is None else k_max
    if k_min > k_max:
        k_min, k_max = k_max, k_min
        warnings.warn("Given k_min larger than k_max. Input values swapped.")
    return k_min, k_max


def _full(shape, noise, unc):
    return np.full(shape, noise, dtype=DTYPE), np.full(shape, unc, dtype=DTYPE)


# ---------------------------------------------------------------- azimuthal average
def _aa_zero(obj: AzimuthalAverage):
    return _full(len(obj.k), 0.0, 0.0)


def _aa_min(obj: AzimuthalAverage):
    i = np.nanargmin(obj.data[:, 0])
    v = obj.data[i, 0]
    return _full(len(obj.k), v, v if obj.err is None else obj.err[i, 0])


def _aa_mean(obj: AzimuthalAverage, values, errs, factor, k_min, k_max):
    k_min, k_max = _k_range(k_min, k_max, np.max(obj.k))
    sel = (obj.k >= k_min) & (obj.k <= k_max)
    v = factor * np.nanmean(values[sel])
    if errs is None:
        u = v
    else:
        n = np.size(~np.isnan(errs[sel]))
        u = factor * np.sqrt(np.nansum(errs[sel] ** 2)) / n if n > 0 else np.nan
    return _full(len(obj.k), v, u)


def _aa_high_q(obj, k_min=None, k_max=None):
    return _aa_mean(obj, obj.data, obj.err, 1, k_min, k_max)


def _aa_power_spec(obj, k_min=None, k_max=None):
    return _aa_mean(obj, obj.power_spec, obj.power_spec_err, 2, k_min, k_max)


def _aa_var(obj, k_min=None, k_max=None):
    return _aa_mean(obj, obj.var, obj.var_err, 2, k_min, k_max)


def _aa_polyfit(obj: AzimuthalAverage, num_points: int = 5):
    deg = 2
    num_points = max(deg + 1, num_points)
    noise, unc = _aa_zero(obj)
    x = obj.tau[:num_points]
    for i in range(len(obj.k)):
        if np.isnan(obj.data[i, 0]):
            noise[i] = unc[i] = np.nan
            continue
        e = None if obj.err is None else obj.err[i, :num_points]
        w = None if e is None or np.any(np.isnan(e)) else 1 / e
        p, cov = np.polyfit(x, obj.data[i, :num_points], deg=deg, w=w, cov="unscaled")
        noise[i], unc[i] = p[-1], np.sqrt(np.diag(cov))[-1]
    return noise, unc


# ------------------------------------------------------- image structure function
def _isf_zero(obj: ImageStructureFunction):
    return _full(obj.shape[1:], 0.0, 0.0)


def _isf_min(obj: ImageStructureFunction, mask=None):
    _, ny, nx = obj.shape
    m = _as_bool_mask(mask, (ny, nx))
    i = np.nanargmin(obj.data[0, m])
    v = obj.data[0, i // nx, i % nx]
    return _full((ny, nx), v, v)


def _isf_select(obj, k_min, k_max, mask):
    k_min, k_max = _k_range(k_min, k_max, max(np.max(obj.kx), np.max(obj.ky)))
    KX, KY = np.meshgrid(obj.kx, obj.ky)
    k = np.sqrt(KX ** 2 + KY ** 2)
    return (k >= k_min) & (k <= k_max) & _as_bool_mask(mask, k.shape)


def _isf_high_q(obj, k_min=None, k_max=None, mask=None):
    v = np.nanmean(obj.data[:, _isf_select(obj, k_min, k_max, mask)])
    return _full(obj.shape[1:], v, v)


def _isf_power_spec(obj, k_min=None, k_max=None, mask=None):
    v = 2 * np.nanmean(obj.power_spec[_isf_select(obj, k_min, k_max, mask)])
    return _full(obj.shape[1:], v, v)


def _isf_var(obj, k_min=None, k_max=None, mask=None):
    v = 2 * np.nanmean(obj.var[_isf_select(obj, k_min, k_max, mask)])
    return _full(obj.shape[1:], v, v)


def _isf_polyfit(obj: ImageStructureFunction, num_points: int = 5):
    deg = 2
    num_points = max(deg + 1, num_points)
    _, ny, nx = obj.shape
    noise, unc = _isf_zero(obj)
    y = obj.data[:num_points].reshape(num_points, -1)
    ok = ~np.isnan(y[0])
    if np.any(ok):
        p, cov = np.polyfit(obj.tau[:num_points], y[:, ok], deg=deg, cov=True)
        noise.reshape(-1)[ok] = p[-1]
        unc.reshape(-1)[ok] = np.sqrt(cov[-1, -1])
    noise.reshape(-1)[~ok] = np.nan
    unc.reshape(-1)[~ok] = np.nan
    return noise, unc


_AA = {"zero": _aa_zero, "min": _aa_min, "high_q": _aa_high_q, "power_spec": _aa_power_spec, "var": _aa_var,
       "polyfit": _aa_polyfit}
_ISF = {"zero": _isf_zero, "min": _isf_min, "high_q": _isf_high_q, "power_spec": _isf_power_spec,
        "var": _isf_var, "polyfit": _isf_polyfit}


def estimate_camera_noise(obj: Union[ImageStructureFunction, AzimuthalAverage], mode: str = "zero",
                          **kwargs) -> Tuple[np.ndarray, np.ndarray]:
    """Noise factor B and its uncertainty, per wavevector (reference
    noise_est.py:759-841).  Raises RuntimeError for an unknown mode and
    TypeError for an unsupported input."""
    if isinstance(obj, ImageStructureFunction):
        table = _ISF
    elif isinstance(obj, AzimuthalAverage):
        table = _AA
    else:
        raise TypeError(f"Input type {type(obj)} not supported.")
    if mode not in table:
        raise RuntimeError(f"Unknown mode '{mode}' selected. Only possible options are {list(table)}.")
    return table[mode](obj=obj, **kwargs)
