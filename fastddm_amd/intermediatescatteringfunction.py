"""Intermediate scattering function estimate from an azimuthal average.

``azavg2isf_estimate`` follows the reference
(src/python/intermediatescatteringfunction.py:733-880): with noise B (from
``estimate_camera_noise`` or given) and plateau A + B (twice VAR or PS, or
given), f(q, dt) = 1 - (D(q, dt) - B) / A, and the propagated uncertainty
sigma^2 = ((D - B) / A^2 sigma_{A+B})^2 + ((A + B - D) / A^2 sigma_B)^2
+ (sigma_D / A)^2.  The container mirrors the reference's
``IntermediateScatteringFunction`` fields (data, err, k, tau, bin_edges).
"""

from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from ._config import DTYPE
from .azimuthalaverage import AzimuthalAverage
from .noise_est import estimate_camera_noise


@dataclass
class IntermediateScatteringFunction:
    _data: np.ndarray
    _err: Optional[np.ndarray]
    k: np.ndarray
    tau: np.ndarray
    bin_edges: np.ndarray

    @property
    def data(self) -> np.ndarray:
        return self._data

    @property
    def err(self) -> Optional[np.ndarray]:
        return self._err

    @property
    def shape(self) -> Tuple[int, int]:
        return self._data.shape


def _custom(values, errs, dim, what):
    if values is not None and len(values) != dim:
        raise RuntimeError(f"Custom {what} array dimension not compatible with given azimuthal average.\n"
                           f"Size of {what} should be {dim}.")
    if errs is None:
        errs = values
    if errs is not None and len(errs) != dim:
        raise RuntimeError(f"Custom {what}_err array dimension not compatible with given azimuthal average.\n"
                           f"Size of {what}_err should be {dim}.")
    return values, errs


def azavg2isf_estimate(az_avg: AzimuthalAverage, noise_est: str = "polyfit", plateau_est: str = "var",
                       noise: Optional[np.ndarray] = None, noise_err: Optional[np.ndarray] = None,
                       plateau: Optional[np.ndarray] = None, plateau_err: Optional[np.ndarray] = None,
                       **kwargs) -> IntermediateScatteringFunction:
    """Convert an ``AzimuthalAverage`` into an ``IntermediateScatteringFunction``."""
    dim_k = az_avg.shape[0]
    if noise_est == "custom":
        noise, noise_err = _custom(noise, noise_err, dim_k, "noise")
    else:
        noise, noise_err = estimate_camera_noise(az_avg, mode=noise_est, **kwargs)
    noise, noise_err = np.asarray(noise).astype(DTYPE), np.asarray(noise_err).astype(DTYPE)

    if plateau_est == "custom":
        plateau, plateau_err = _custom(plateau, plateau_err, dim_k, "plateau")
    elif plateau_est in ("var", "power_spec"):
        src = az_avg.var if plateau_est == "var" else az_avg.power_spec
        src_err = az_avg.var_err if plateau_est == "var" else az_avg.power_spec_err
        plateau = 2 * src
        plateau_err = 2 * (src if src_err is None else src_err)
    else:
        raise RuntimeError(f"Unsupported plateau_est mode {plateau_est}. "
                           f"Possible values are {['custom', 'power_spec', 'var']}")
    plateau, plateau_err = np.asarray(plateau).astype(DTYPE), np.asarray(plateau_err).astype(DTYPE)

    y = az_avg.data.astype(DTYPE)
    sy = y if az_avg.err is None else az_avg.err.astype(DTYPE)
    ApB, B = plateau[:, None], noise[:, None]
    A = ApB - B
    s_ApB, s_B = plateau_err[:, None], noise_err[:, None]
    data = 1 - (y - B) / A
    err = np.sqrt(((y - B) / A ** 2 * s_ApB) ** 2 + ((ApB - y) / A ** 2 * s_B) ** 2 + (sy / A) ** 2)
    return IntermediateScatteringFunction(data.astype(DTYPE), err.astype(DTYPE), az_avg.k.astype(DTYPE),
                                          np.asarray(az_avg.tau).astype(DTYPE), az_avg.bin_edges.astype(DTYPE))
