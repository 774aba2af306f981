"""The ``'hip'`` backend core: differential dynamic microscopy on one AMD GPU.

Same positional contract as the reference's CUDA adapters
(src/python/_ddm_cuda.py:17-113): ``(img_seq, lags, nx, ny[, nt], window,
gpu_id=0, **kwargs) -> ndarray (NL+2, ny, nx//2+1)`` of ``DTYPE``, already
fftshifted along y.  ``set_device(gpu_id)`` is called first, as there.

Extra keyword: ``precision`` in {'fp64', 'fp32', 'fp32c'} (default 'fp64',
the precision of the reference GPU core).  Other keywords are ignored, as the
reference's ``**kwargs`` are.
"""

from typing import List, Optional

import numpy as np

from . import _core_hip
from ._config import DTYPE


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
    _core_hip.set_device(gpu_id)
    return _core_hip.ddm_fft(img_seq, lags, nx, ny, nt, window, precision=precision, out_dtype=DTYPE)
