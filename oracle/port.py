"""TEST INFRASTRUCTURE ONLY -- ctypes front of the C restatement (liboracle.so).

Mirrors the reference's ``_core.ddm_fft`` / ``_core.ddm_diff`` call shapes
(src/python/bindings/ddm_bindings.h:24-43): image stack, lags, nx, ny,
(nt, chunk_size,) window -> (NL+2, ny, nx//2+1) float64.
"""

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

DTYPE_CODES = {
    np.dtype(np.uint8): 0,
    np.dtype(np.int16): 1,
    np.dtype(np.uint16): 2,
    np.dtype(np.int32): 3,
    np.dtype(np.uint32): 4,
    np.dtype(np.int64): 5,
    np.dtype(np.uint64): 6,
    np.dtype(np.float32): 7,
    np.dtype(np.float64): 8,
}

_lib = None


def build():
    """Compile liboracle.so with the Makefile beside this file."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.isfile(_LIB_PATH):
            raise RuntimeError(f"{_LIB_PATH} is not built (make -C oracle)")
        L = ctypes.CDLL(_LIB_PATH)
        u64, p, i32 = ctypes.c_uint64, ctypes.c_void_p, ctypes.c_int
        L.oracle_ddm_fft.argtypes = [p, i32, u64, u64, u64, p, u64, u64, u64, u64, u64, p, p, i32]
        L.oracle_ddm_diff.argtypes = [p, i32, u64, u64, u64, p, u64, u64, u64, p, p, i32]
        L.oracle_rfft2.argtypes = [p, i32, u64, u64, u64, u64, u64, p, p, i32]
        L.oracle_fft.argtypes = [p, u64, u64]
        for f in (L.oracle_ddm_fft, L.oracle_ddm_diff, L.oracle_rfft2, L.oracle_fft):
            f.restype = i32
        _lib = L
    return _lib


def _prep(img, lags, window):
    img = np.ascontiguousarray(img)
    if img.dtype not in DTYPE_CODES:
        raise TypeError(f"unsupported dtype {img.dtype}")
    lags = np.ascontiguousarray(np.asarray(lags, dtype=np.uint32))
    w = None
    if window is not None and np.size(window) > 0:
        w = np.ascontiguousarray(window, dtype=np.float64)
    return img, lags, w


def ddm_fft(img, lags, nx, ny, nt, window=None, chunk_size=None, threads=1):
    img, lags, w = _prep(img, lags, window)
    L, H, W = img.shape
    nq = ny * (nx // 2 + 1)
    chunk = int(chunk_size or min(nq, 256))
    out = np.empty((len(lags) + 2, ny, nx // 2 + 1), dtype=np.float64)
    rc = lib().oracle_ddm_fft(
        img.ctypes.data, DTYPE_CODES[img.dtype], L, H, W, lags.ctypes.data, len(lags),
        nx, ny, nt, chunk, None if w is None else w.ctypes.data, out.ctypes.data, threads,
    )
    if rc:
        raise RuntimeError(f"oracle_ddm_fft failed ({rc})")
    return out


def ddm_diff(img, lags, nx, ny, window=None, threads=1):
    img, lags, w = _prep(img, lags, window)
    L, H, W = img.shape
    out = np.empty((len(lags) + 2, ny, nx // 2 + 1), dtype=np.float64)
    rc = lib().oracle_ddm_diff(
        img.ctypes.data, DTYPE_CODES[img.dtype], L, H, W, lags.ctypes.data, len(lags),
        nx, ny, None if w is None else w.ctypes.data, out.ctypes.data, threads,
    )
    if rc:
        raise RuntimeError(f"oracle_ddm_diff failed ({rc})")
    return out


def rfft2(img, nx, ny, window=None, threads=1):
    img, _, w = _prep(img, [], window)
    L, H, W = img.shape
    out = np.empty((L, ny, nx // 2 + 1), dtype=np.complex128)
    lib().oracle_rfft2(
        img.ctypes.data, DTYPE_CODES[img.dtype], L, H, W, nx, ny,
        None if w is None else w.ctypes.data, out.ctypes.data, threads,
    )
    return out


def fft(x):
    """Forward FFT along the last axis (complex128), via the C Stockham FFT."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = x.shape[-1]
    lib().oracle_fft(x.ctypes.data, n, x.size // n)
    return x
