"""Build and precision flags of the package.

Mirrors the reference's generated ``_config.py`` (src/python/_config.py.in:10-16):
``DTYPE`` is float64 unless single precision is selected, in which case it is
float32. The reference fixes this at CMake time (``SINGLE_PRECISION``); here it
is read once from the environment variable ``FASTDDM_SINGLE_PRECISION`` at
import time, so the same installed tree serves both builds.

``IS_HIP_ENABLED`` plays the role of ``IS_CPP_ENABLED``/``IS_CUDA_ENABLED``:
the ``'hip'`` core is registered only when the compiled HIP library sits next
to this package (``libfddm_hip.so``).
"""

import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIBRARY_PATH = os.path.join(_HERE, "libfddm_hip.so")

IS_CPP_ENABLED = False
IS_CUDA_ENABLED = False
IS_HIP_ENABLED = os.path.isfile(HIP_LIBRARY_PATH)
IS_SINGLE_PRECISION = os.environ.get("FASTDDM_SINGLE_PRECISION", "0").lower() in (
    "1",
    "on",
    "true",
    "yes",
)

DTYPE = np.float32 if IS_SINGLE_PRECISION else np.float64

__all__ = [
    "IS_CPP_ENABLED",
    "IS_CUDA_ENABLED",
    "IS_HIP_ENABLED",
    "IS_SINGLE_PRECISION",
    "DTYPE",
]
