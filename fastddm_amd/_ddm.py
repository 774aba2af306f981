r"""Differential dynamic microscopy front end with a backend registry.

The image structure function of an image sequence I(x, t) is

.. math::

    D(\vec{q},\Delta t)= \langle \lvert \tilde{I}(\vec{q},t+\Delta t) -
    \tilde{I}(\vec{q},t) \rvert^2 \rangle_t

with the power spectrum PS(q) = <|I~(q, t)|^2>_t and the variance over time
VAR(q) = PS(q) - |<I~(q, t)>_t|^2.

``ddm()`` keeps the reference's contract (src/python/_ddm.py:77-199): the same
argument checks and messages, lag sanitising, padded sizes, backend call
signature and ``ImageStructureFunction`` result. The registry
(reference ``_backend``, ``_ddm.py:51-74``) holds the ``'hip'`` core when the
HIP library is built; the reference's CPU cores are not part of this package.
"""

from typing import Callable, Dict, Iterable, Optional
import warnings

import numpy as np

from ._config import DTYPE, IS_HIP_ENABLED
from ._fftopt import next_fast_len
from .imagestructurefunction import ImageStructureFunction

_backend: Dict[str, Dict[str, Callable]] = {}

if IS_HIP_ENABLED:
    from ._ddm_hip import ddm_diff_hip, ddm_fft_hip

    _backend["hip"] = {"diff": ddm_diff_hip, "fft": ddm_fft_hip}


def _prepare_lags(lags: Iterable[int], length: int) -> np.ndarray:
    """Sorted unique positive lags, as ``_ddm.py:138-159`` of the reference.

    An empty iterable fails in ``np.min`` (ValueError) exactly as in the
    reference; ``[0]`` alone leaves no lags, and the result then holds only
    the power spectrum and the variance.
    """
    lags = np.array(sorted(set(lags)), dtype=np.int64)
    if np.min(lags) < 0:
        raise RuntimeError("Negative lags are not possible.")
    if np.max(lags) > length - 1:
        raise RuntimeError(
            f"Lags larger than len(img_seq) - 1 = {length - 1} are not possible."
        )
    if lags[0] == 0:
        warnings.warn("Found 0 in lags. Removed for compatibility with other functions.")
        lags = lags[1:]
    return lags


def ddm(
    img_seq: np.ndarray,
    lags: Iterable[int],
    *,
    core: str = "hip",
    mode: str = "fft",
    window: Optional[np.ndarray] = None,
    **kwargs,
) -> ImageStructureFunction:
    """Differential Dynamic Microscopy analysis of an image sequence.

    Parameters
    ----------
    img_seq : numpy.ndarray
        Image sequence of shape ``(t, y, x)``; any of uint8, int16, uint16,
        int32, uint32, int64, uint64, float32, float64.
    lags : Iterable[int]
        Delays (in frames) to evaluate.
    core : str, optional
        Backend core. Default is ``"hip"``.
    mode : str, optional
        ``"fft"`` (Wiener-Khinchin) or ``"diff"`` (direct differences).
    window : numpy.ndarray, optional
        2D window of shape ``(y, x)`` applied to every image.
    **kwargs
        Forwarded to the backend: ``gpu_id`` (default 0) and
        ``precision`` in {``'fp64'``, ``'fp32'``, ``'fp32c'``} for ``'hip'``.

    Raises
    ------
    RuntimeError
        On an unknown core or mode, negative or too large lags, or a window
        whose shape differs from the images.
    """
    backend = _backend
    cores = list(backend.keys())
    modes = ["diff", "fft"]

    if core not in cores:
        raise RuntimeError(
            f"Unknown core '{core}' selected. Only possible options are {cores}."
        )
    if mode not in modes:
        raise RuntimeError(
            f"Unknown mode '{mode}' selected. Only possible options are {modes}."
        )

    dim_t, dim_y, dim_x = img_seq.shape
    lags = _prepare_lags(lags, dim_t)

    if window is None:
        window = np.array([], dtype=DTYPE)
    elif window.shape != (dim_y, dim_x):
        raise RuntimeError(
            f"Window with shape {window.shape} incompatible with image shape {(dim_y, dim_x)}."
        )

    dim_x_padded = next_fast_len(dim_x, core)
    dim_y_padded = next_fast_len(dim_y, core)

    ddm_func = backend[core][mode]
    args = [img_seq, lags, dim_x_padded, dim_y_padded]
    if mode == "fft":
        args.append(next_fast_len(2 * dim_t, core))
    args.append(window.astype(DTYPE))

    kx = 2 * np.pi * np.fft.fftfreq(dim_x_padded)[: (dim_x_padded // 2 + 1)]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(dim_y_padded))
    return ImageStructureFunction(
        ddm_func(*args, **kwargs),
        kx.astype(DTYPE),
        ky.astype(DTYPE),
        dim_x_padded,
        dim_y_padded,
        lags.astype(DTYPE),
    )
