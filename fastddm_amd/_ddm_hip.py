"""The ``'hip'`` backend core: differential dynamic microscopy on one AMD GPU.

Same positional contract as the reference's CUDA adapters
(src/python/_ddm_cuda.py:17-113): ``(img_seq, lags, nx, ny[, nt], window,
gpu_id=0, **kwargs) -> ndarray (NL+2, ny, nx//2+1)`` of ``DTYPE``, already
fftshifted along y.  ``set_device(gpu_id)`` is called first, as there.

Extra keyword: ``precision`` in {'fp64', 'fp32', 'fp32c'} (default 'fp64',
the precision of the reference GPU core).  Other keywords are ignored, as the
reference's ``**kwargs`` are.

'fp32' (float spectra, sums and differences with no compensation) is kept for
precision studies only: near q = 0, where |F| is dominated by the mean
intensity, D cancels catastrophically in float and misses the reference's
single-precision tolerance (rtol 1e-3, tests/test_ddm.py:116-157) by orders of
magnitude, so selecting it warns.  'fp32c' meets that tolerance.
"""

import warnings
from typing import List, Optional

import numpy as np

from . import _core_hip
from ._config import DTYPE


def _warn_plain_fp32(precision: str) -> None:
    if precision == "fp32":
        warnings.warn(
            "precision='fp32' does not meet the single-precision tolerance (rtol 1e-3) near q = 0, "
            "where D cancels in float; use precision='fp32c' (compensated) or 'fp64'.",
            RuntimeWarning,
            stacklevel=3,
        )


def ddm_diff_hip(
    img_seq: np.ndarray,
    lags: List[int],
    nx: int,
    ny: int,
    window: np.ndarray,
    gpu_id: Optional[int] = 0,
    precision: str = "fp64",
    **kwargs,
) -> np.ndarray:
    """diff mode: D(q, dt) = <|F(q, t + dt) - F(q, t)|^2>_t on the GPU.

    Raises
    ------
    RuntimeError
        On a bad device id, unsupported sizes, or insufficient device memory.
    """
    _warn_plain_fp32(precision)
    _core_hip.set_device(gpu_id)
    return _core_hip.ddm_diff(img_seq, lags, nx, ny, window, precision=precision, out_dtype=DTYPE)


def ddm_fft_hip(
    img_seq: np.ndarray,
    lags: List[int],
    nx: int,
    ny: int,
    nt: int,
    window: np.ndarray,
    gpu_id: Optional[int] = 0,
    precision: str = "fp64",
    **kwargs,
) -> np.ndarray:
    """fft mode: Wiener-Khinchin form of D(q, dt) on the GPU.

    Raises
    ------
    RuntimeError
        On a bad device id, unsupported sizes, or insufficient device memory.
    """
    _warn_plain_fp32(precision)
    _core_hip.set_device(gpu_id)
    return _core_hip.ddm_fft(img_seq, lags, nx, ny, nt, window, precision=precision, out_dtype=DTYPE)
