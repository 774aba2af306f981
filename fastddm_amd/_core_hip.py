"""ctypes binding of ``libfddm_hip.so`` (declared in include/fddm_hip.h).

Plays the role of the reference's pybind11 module ``_core_cuda``
(src/python/module_cuda.cc, src/python/bindings/ddm_cuda_bindings.h:24-56):
device utilities plus the fft/diff entry points.  Every failing call raises
``RuntimeError`` with the library's message; the library is required -- there
is no fallback path.  ctypes releases the GIL for the duration of each call.
"""

import ctypes
from typing import Optional

import numpy as np

from ._config import HIP_LIBRARY_PATH

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
PRECISIONS = {"fp64": 0, "fp32": 1, "fp32c": 2}
OUT_CODES = {np.dtype(np.float64): 0, np.dtype(np.float32): 1}
MODE_DIFF, MODE_FFT = 0, 1
_ERRLEN = 1024

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        L = ctypes.CDLL(HIP_LIBRARY_PATH)
    except OSError as exc:
        raise RuntimeError(
            f"fastddm_amd HIP library not loadable ({HIP_LIBRARY_PATH}): {exc}. "
            "Build it with `python -m fastddm_amd._build`."
        ) from exc
    p, u64, i32, sz = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
    cp = ctypes.c_char_p
    sig = {
        "fddm_get_num_devices": [ctypes.POINTER(i32), cp, sz],
        "fddm_set_device": [i32, cp, sz],
        "fddm_get_device": [ctypes.POINTER(i32), cp, sz],
        "fddm_get_free_device_memory": [ctypes.POINTER(u64), cp, sz],
        "fddm_get_free_host_memory": [ctypes.POINTER(u64), cp, sz],
        "fddm_get_host_memory": [i32, u64, u64, u64, u64, i32, ctypes.POINTER(u64), cp, sz],
        "fddm_check_host_memory": [i32, u64, u64, u64, u64, i32, ctypes.POINTER(i32), cp, sz],
        "fddm_ddm_fft": [p, i32, u64, u64, u64, p, u64, u64, u64, u64, p, i32, i32, i32, p, cp, sz],
        "fddm_ddm_diff": [p, i32, u64, u64, u64, p, u64, u64, u64, p, i32, i32, i32, p, cp, sz],
        "fddm_release_cache": [cp, sz],
        "fddm_set_memory_budget": [u64, cp, sz],
        "fddm_set_cache_limit": [ctypes.c_int64, cp, sz],
        "fddm_get_cached_bytes": [p, cp, sz],
        "fddm_get_last_call_chunks": [ctypes.POINTER(u64), cp, sz],
        "fddm_ingest_begin": [ctypes.POINTER(p), i32, i32, u64, u64, u64, p, u64, u64, u64, u64, p, i32, i32, i32,
                              cp, sz],
        "fddm_ingest_push": [p, p, u64, cp, sz],
        "fddm_ingest_finish": [p, p, cp, sz],
        "fddm_ingest_destroy": [p],
        "fddm_plan_create": [ctypes.POINTER(p), i32, i32, u64, u64, u64, p, u64, u64, u64, u64,
                             p, i32, i32, i32, cp, sz],
        "fddm_plan_destroy": [p],
        "fddm_plan_workspace_size": [p, ctypes.POINTER(u64)],
        "fddm_plan_output_size": [p, ctypes.POINTER(u64)],
        "fddm_plan_execute": [p, p, p, p, p, cp, sz],
        "fddm_plan_execute_static": [p, p, p, p, cp, sz],
        "fddm_plan_timing_begin": [p, i32, cp, sz],
        "fddm_plan_timing_end": [p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(i32), cp, sz],
        "fddm_plan_spectra": [p, p, p, p, cp, sz],
        "fddm_malloc": [ctypes.POINTER(p), u64, cp, sz],
        "fddm_free": [p],
        "fddm_memcpy_htod": [p, p, u64, cp, sz],
        "fddm_memcpy_dtoh": [p, p, u64, cp, sz],
        "fddm_memcpy_dtod": [p, p, u64, cp, sz],
        "fddm_stream_create": [ctypes.POINTER(p), cp, sz],
        "fddm_stream_destroy": [p],
        "fddm_stream_synchronize": [p, cp, sz],
        "fddm_test_fft": [p, u64, u64, i32, cp, sz],
        "fddm_server_create": [ctypes.POINTER(p), i32, u64, cp, sz],
        "fddm_server_destroy": [p],
        "fddm_server_set_aging": [p, ctypes.c_double],
        "fddm_server_submit": [p, i32, p, i32, i32, u64, u64, u64, p, u64, u64, u64, u64, p, i32,
                               i32, i32, ctypes.POINTER(u64), cp, sz],
        "fddm_server_status": [p, u64, ctypes.POINTER(i32), ctypes.POINTER(ctypes.c_double), cp, sz],
        "fddm_server_wait": [p, u64, ctypes.c_double, cp, sz],
        "fddm_server_next_done": [p, ctypes.c_double, ctypes.POINTER(u64), cp, sz],
        "fddm_server_result": [p, u64, ctypes.POINTER(p), ctypes.POINTER(u64), cp, sz],
        "fddm_server_fetch": [p, u64, p, cp, sz],
        "fddm_server_release": [p, u64],
        "fddm_server_stats": [p, ctypes.POINTER(u64)],
        "fddm_fill_uniform": [p, u64, i32, u64, ctypes.c_uint32, ctypes.c_uint32, p, cp, sz],
        "fddm_batch_create": [ctypes.POINTER(p), p, i32, i32, cp, sz],
        "fddm_batch_launch": [p, p, cp, sz],
        "fddm_batch_launch_eager": [p, p, cp, sz],
        "fddm_batch_frames_per_block": [p, ctypes.POINTER(i32)],
        "fddm_batch_spatial_kernels": [p, ctypes.POINTER(i32), ctypes.POINTER(i32)],
        "fddm_batch_destroy": [p],
        "fddm_azimuthal_average": [p, i32, u64, u64, u64, p, p, p, i32, i32, p, p, p, cp, sz],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = i32
    _lib = L
    return L


EXPORTED_SYMBOLS = (
    "fddm_get_num_devices",
    "fddm_set_device",
    "fddm_get_device",
    "fddm_get_free_device_memory",
    "fddm_get_free_host_memory",
    "fddm_get_host_memory",
    "fddm_check_host_memory",
    "fddm_ddm_fft",
    "fddm_ddm_diff",
    "fddm_release_cache",
    "fddm_set_memory_budget",
    "fddm_set_cache_limit",
    "fddm_get_cached_bytes",
    "fddm_get_last_call_chunks",
    "fddm_ingest_begin",
    "fddm_ingest_push",
    "fddm_ingest_finish",
    "fddm_ingest_destroy",
    "fddm_plan_create",
    "fddm_plan_destroy",
    "fddm_plan_workspace_size",
    "fddm_plan_output_size",
    "fddm_plan_execute",
    "fddm_plan_execute_static",
    "fddm_plan_timing_begin",
    "fddm_plan_timing_end",
    "fddm_plan_spectra",
    "fddm_malloc",
    "fddm_free",
    "fddm_memcpy_htod",
    "fddm_memcpy_dtoh",
    "fddm_memcpy_dtod",
    "fddm_stream_create",
    "fddm_stream_destroy",
    "fddm_stream_synchronize",
    "fddm_test_fft",
    "fddm_server_create",
    "fddm_server_destroy",
    "fddm_server_set_aging",
    "fddm_server_submit",
    "fddm_server_status",
    "fddm_server_wait",
    "fddm_server_next_done",
    "fddm_server_result",
    "fddm_server_fetch",
    "fddm_server_release",
    "fddm_server_stats",
    "fddm_fill_uniform",
    "fddm_batch_create",
    "fddm_batch_launch",
    "fddm_batch_launch_eager",
    "fddm_batch_frames_per_block",
    "fddm_batch_spatial_kernels",
    "fddm_batch_destroy",
    "fddm_azimuthal_average",
)


def _call(fname: str, *args) -> None:
    err = ctypes.create_string_buffer(_ERRLEN)
    rc = getattr(lib(), fname)(*args, err, _ERRLEN)
    if rc != 0:
        msg = err.value.decode(errors="replace") or f"error code {rc}"
        raise RuntimeError(f"{fname}: {msg}")


# ---- device utilities (reference ddm_cuda_bindings.h:26-31) ----
def get_num_devices() -> int:
    n = ctypes.c_int(0)
    _call("fddm_get_num_devices", ctypes.byref(n))
    return n.value


def set_device(gpu_id: int) -> None:
    _call("fddm_set_device", int(gpu_id))


def get_device() -> int:
    d = ctypes.c_int(0)
    _call("fddm_get_device", ctypes.byref(d))
    return d.value


def get_free_device_memory() -> int:
    f = ctypes.c_uint64(0)
    _call("fddm_get_free_device_memory", ctypes.byref(f))
    return f.value


def get_free_host_memory() -> int:
    """Available host memory in bytes (reference get_free_host_memory,
    src/cuda/memchk_gpu.cc:35-87)."""
    f = ctypes.c_uint64(0)
    _call("fddm_get_free_host_memory", ctypes.byref(f))
    return f.value


def _host_memory(mode: int, nx, ny, length, num_lags, out_dtype) -> int:
    b = ctypes.c_uint64(0)
    _call("fddm_get_host_memory", mode, nx, ny, length, num_lags, _out_code(out_dtype), ctypes.byref(b))
    return b.value


def get_host_memory_fft(nx, ny, length, num_lags, out_dtype=np.float64) -> int:
    """Host bytes an fft-mode call needs besides its input: the result and
    the pinned staging buffers (reference get_host_memory_fft,
    memchk_gpu.cc:116-140)."""
    return _host_memory(MODE_FFT, nx, ny, length, num_lags, out_dtype)


def get_host_memory_diff(nx, ny, length, num_lags, out_dtype=np.float64) -> int:
    """As get_host_memory_fft, for diff mode (memchk_gpu.cc:89-114)."""
    return _host_memory(MODE_DIFF, nx, ny, length, num_lags, out_dtype)


def _check_host_memory(mode: int, nx, ny, length, num_lags, out_dtype) -> bool:
    ok = ctypes.c_int(0)
    _call("fddm_check_host_memory", mode, nx, ny, length, num_lags, _out_code(out_dtype), ctypes.byref(ok))
    return bool(ok.value)


def check_host_memory_fft(nx, ny, length, num_lags, out_dtype=np.float64) -> bool:
    """True when 0.95 x the available host memory exceeds that need
    (reference check_host_memory_fft_py, memchk_gpu.cc:172-190)."""
    return _check_host_memory(MODE_FFT, nx, ny, length, num_lags, out_dtype)


def check_host_memory_diff(nx, ny, length, num_lags, out_dtype=np.float64) -> bool:
    """As check_host_memory_fft, for diff mode (memchk_gpu.cc:142-160)."""
    return _check_host_memory(MODE_DIFF, nx, ny, length, num_lags, out_dtype)


def release_cache() -> None:
    _call("fddm_release_cache")


def set_cache_limit(nbytes: int) -> None:
    """Device bytes of buffers ``ddm_fft`` / ``ddm_diff`` keep on the current
    device for the next call (negative: the default, 16 GiB; 0: none).  A
    call needing more frees its buffers before returning."""
    _call("fddm_set_cache_limit", int(nbytes))


def cached_bytes() -> int:
    """Device bytes the host-path calls keep cached on the current device."""
    v = ctypes.c_uint64(0)
    _call("fddm_get_cached_bytes", ctypes.byref(v))
    return v.value


def set_memory_budget(nbytes: int) -> None:
    """Cap the device memory of ``ddm_fft`` / ``ddm_diff`` on the current
    device (0 restores the default, 95 % of free memory).  Larger analyses
    then run in chunks of frames and of q, with the same result."""
    _call("fddm_set_memory_budget", int(nbytes))


def last_call_chunks() -> dict:
    """How the last ``ddm_fft`` / ``ddm_diff`` call on the current device ran:
    mode 0 (one launch per stage), 1 (pieces, spectra on the device) or 2
    (pieces, spectra in host memory), and its frame and q chunk counts."""
    v = (ctypes.c_uint64 * 3)()
    _call("fddm_get_last_call_chunks", v)
    return {"mode": int(v[0]), "frame_chunks": int(v[1]), "q_chunks": int(v[2])}


# ---- argument preparation ----
def _img_arg(img_seq: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img_seq)
    if img.ndim != 3:
        raise RuntimeError(f"img_seq must be 3D (t, y, x), got shape {img.shape}")
    if img.dtype not in DTYPE_CODES:
        raise RuntimeError(f"unsupported image dtype {img.dtype}")
    return img


def _lags_arg(lags) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(lags, dtype=np.uint32))


def _window_arg(window: Optional[np.ndarray], height: int, width: int):
    """(pointer, is_float32, keep-alive array) of a window of shape (height, width).

    The library reads height * width values from the pointer, so any other
    shape is refused here (message of the reference's check, _ddm.py:161-168).
    """
    if window is None or np.size(window) == 0:
        return None, 0, None
    w = np.ascontiguousarray(window)
    if w.shape != (int(height), int(width)):
        raise RuntimeError(
            f"Window with shape {w.shape} incompatible with image shape {(int(height), int(width))}."
        )
    if w.dtype == np.float32:
        return w.ctypes.data, 1, w
    w = np.ascontiguousarray(w, dtype=np.float64)
    return w.ctypes.data, 0, w


def _out_code(out_dtype) -> int:
    dt = np.dtype(out_dtype)
    if dt not in OUT_CODES:
        raise RuntimeError(f"unsupported output dtype {dt}")
    return OUT_CODES[dt]


def _prec_code(precision: str) -> int:
    if precision not in PRECISIONS:
        raise RuntimeError(
            f"Unknown precision '{precision}'. Only possible options are {list(PRECISIONS)}."
        )
    return PRECISIONS[precision]


def ddm_fft(img_seq, lags, nx, ny, nt, window, precision="fp64", out_dtype=np.float64):
    """fft mode on the current device; returns (NL+2, ny, nx//2+1) of out_dtype."""
    img = _img_arg(img_seq)
    lg = _lags_arg(lags)
    L, H, W = img.shape
    wp, wf32, _keep = _window_arg(window, H, W)
    if not check_host_memory_fft(nx, ny, L, len(lg), out_dtype):
        raise RuntimeError("Not enough space in memory to store the result.\n")
    out = np.empty((len(lg) + 2, ny, nx // 2 + 1), dtype=out_dtype)
    _call(
        "fddm_ddm_fft", img.ctypes.data, DTYPE_CODES[img.dtype], L, H, W, lg.ctypes.data, len(lg),
        nx, ny, nt, wp, wf32, _prec_code(precision), _out_code(out_dtype), out.ctypes.data,
    )
    return out


def ddm_diff(img_seq, lags, nx, ny, window, precision="fp64", out_dtype=np.float64):
    """diff mode on the current device; returns (NL+2, ny, nx//2+1) of out_dtype."""
    img = _img_arg(img_seq)
    lg = _lags_arg(lags)
    L, H, W = img.shape
    wp, wf32, _keep = _window_arg(window, H, W)
    if not check_host_memory_diff(nx, ny, L, len(lg), out_dtype):
        raise RuntimeError("Not enough space in memory to store the result.\n")
    out = np.empty((len(lg) + 2, ny, nx // 2 + 1), dtype=out_dtype)
    _call(
        "fddm_ddm_diff", img.ctypes.data, DTYPE_CODES[img.dtype], L, H, W, lg.ctypes.data, len(lg),
        nx, ny, wp, wf32, _prec_code(precision), _out_code(out_dtype), out.ctypes.data,
    )
    return out


def device_fft(x: np.ndarray, precision: str = "fp64") -> np.ndarray:
    """Forward FFT along the last axis with the device FFT engine (test entry)."""
    a = np.ascontiguousarray(x, dtype=np.complex128).copy()
    n = a.shape[-1]
    _call("fddm_test_fft", a.ctypes.data, n, a.size // n, _prec_code(precision))
    return a


# ---- device buffers, streams and plans ----
class DeviceBuffer:
    """Device allocation owned by Python; freed on ``free()`` or collection."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        _call("fddm_malloc", ctypes.byref(self.ptr), self.nbytes)

    @classmethod
    def from_array(cls, arr: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        buf = cls(a.nbytes)
        _call("fddm_memcpy_htod", buf.ptr, a.ctypes.data, a.nbytes)
        return buf

    def to_array(self, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise RuntimeError("device buffer smaller than requested array")
        _call("fddm_memcpy_dtoh", out.ctypes.data, self.ptr, out.nbytes)
        return out

    def read(self, offset: int, shape, dtype) -> np.ndarray:
        """Copy ``prod(shape)`` elements of ``dtype`` starting ``offset`` bytes in."""
        out = np.empty(shape, dtype=dtype)
        if offset < 0 or offset + out.nbytes > self.nbytes:
            raise RuntimeError("read outside the device buffer")
        _call("fddm_memcpy_dtoh", out.ctypes.data, ctypes.c_void_p(self.ptr.value + int(offset)), out.nbytes)
        return out

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        """Device-to-device copy of nbytes from `other` into this buffer."""
        if nbytes < 0 or nbytes > self.nbytes or nbytes > other.nbytes:
            raise RuntimeError(f"copy of {nbytes} bytes between buffers of {other.nbytes} and {self.nbytes} bytes")
        _call("fddm_memcpy_dtod", self.ptr, other.ptr, int(nbytes))

    def free(self) -> None:
        if self.ptr is not None and self.ptr.value:
            lib().fddm_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        self.handle = ctypes.c_void_p()
        _call("fddm_stream_create", ctypes.byref(self.handle))

    def synchronize(self) -> None:
        _call("fddm_stream_synchronize", self.handle)

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            lib().fddm_stream_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


class Ingest:
    """Streamed analysis of one stack whose frames arrive in pieces.

    ``push`` takes (k, H, W) frames in stack order (any k); their transfer and
    row / column transforms overlap later pushes.  ``finish`` returns the
    (NL+2, ny, nx//2+1) result once all L frames are pushed -- the values of
    ``ddm_fft`` / ``ddm_diff`` on the whole stack.
    """

    def __init__(self, mode, dtype, shape, lags, nx, ny, nt=0, window=None, precision="fp64",
                 out_dtype=np.float64):
        lg = _lags_arg(lags)
        L, H, W = shape
        wp, wf32, _keep = _window_arg(window, H, W)
        self.shape = tuple(int(v) for v in shape)
        self.dtype = np.dtype(dtype)
        self.out_shape = (len(lg) + 2, ny, nx // 2 + 1)
        self.out_dtype = np.dtype(out_dtype)
        self._h = ctypes.c_void_p()
        _call("fddm_ingest_begin", ctypes.byref(self._h), MODE_FFT if mode == "fft" else MODE_DIFF,
              DTYPE_CODES[self.dtype], L, H, W, lg.ctypes.data, len(lg), nx, ny, nt, wp, wf32,
              _prec_code(precision), _out_code(out_dtype))

    def push(self, frames: np.ndarray) -> None:
        a = np.ascontiguousarray(frames, dtype=self.dtype)
        if a.ndim == 2:
            a = a[None]
        if a.shape[1:] != self.shape[1:]:
            raise RuntimeError(f"frames of shape {a.shape[1:]} pushed into a {self.shape[1:]} stack")
        _call("fddm_ingest_push", self._h, a.ctypes.data, a.shape[0])

    def finish(self, out: np.ndarray = None) -> np.ndarray:
        """The result, written into `out` when given (C-contiguous, of the
        result's shape and dtype), else into a new array."""
        if out is None:
            out = np.empty(self.out_shape, dtype=self.out_dtype)
        elif (out.shape != self.out_shape or out.dtype != self.out_dtype or not out.flags.c_contiguous
              or not out.flags.writeable):
            raise RuntimeError(f"output array must be a writeable C-contiguous {self.out_shape} {self.out_dtype} array")
        _call("fddm_ingest_finish", self._h, out.ctypes.data)
        return out

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fddm_ingest_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Plan:
    """A reusable execution plan for device-resident stacks of one shape."""

    STAGES = ("frame0_spectra", "row_pass", "column_pass", "time_or_diff", "edge_fp64")

    def __init__(self, mode, dtype, shape, lags, nx, ny, nt=0, window=None,
                 precision="fp64", out_dtype=np.float64):
        self.handle = ctypes.c_void_p()
        lg = _lags_arg(lags)
        L, H, W = shape
        wp, wf32, _keep = _window_arg(window, H, W)
        self.mode = mode
        self.shape = tuple(shape)
        self.dtype = np.dtype(dtype)
        self.nl = len(lg)
        self.nx, self.ny, self.nt = nx, ny, nt
        self.out_dtype = np.dtype(out_dtype)
        _call(
            "fddm_plan_create", ctypes.byref(self.handle), MODE_FFT if mode == "fft" else MODE_DIFF,
            DTYPE_CODES[np.dtype(dtype)], L, H, W, lg.ctypes.data, len(lg), nx, ny, nt, wp, wf32,
            _prec_code(precision), _out_code(out_dtype),
        )

    @property
    def out_shape(self):
        return (self.nl + 2, self.ny, self.nx // 2 + 1)

    def workspace_size(self) -> int:
        b = ctypes.c_uint64(0)
        lib().fddm_plan_workspace_size(self.handle, ctypes.byref(b))
        return b.value

    def output_size(self) -> int:
        b = ctypes.c_uint64(0)
        lib().fddm_plan_output_size(self.handle, ctypes.byref(b))
        return b.value

    def input_size(self) -> int:
        """Bytes of the (L, H, W) input stack."""
        L, H, W = self.shape
        return L * H * W * self.dtype.itemsize

    def _check_buffers(self, d_img, d_ws, d_out=None) -> None:
        # the kernels index these buffers by the plan's sizes: a short buffer
        # would be read or written out of bounds on the device
        for name, buf, need in (("input", d_img, self.input_size()),
                                ("workspace", d_ws, self.workspace_size()),
                                ("output", d_out, self.output_size() if d_out is not None else 0)):
            if buf is None:
                continue
            if buf.ptr is None or not buf.ptr.value:
                raise RuntimeError(f"{name} buffer has been freed")
            if buf.nbytes < need:
                raise RuntimeError(f"{name} buffer holds {buf.nbytes} bytes, the plan needs {need}")

    def execute(self, d_img: DeviceBuffer, d_out: DeviceBuffer, d_ws: DeviceBuffer,
                stream: Optional[Stream] = None) -> None:
        self._check_buffers(d_img, d_ws, d_out)
        _call("fddm_plan_execute", self.handle, d_img.ptr, d_out.ptr, d_ws.ptr,
              stream.handle if stream is not None else None)

    STATIC_SCRATCH_BYTES = 256 << 20

    def execute_static(self, d_img: DeviceBuffer, d_out: DeviceBuffer, stream: Optional[Stream] = None) -> None:
        """``execute`` with the library's static per-device workspace
        (256 MiB; no allocation): plans with a larger workspace are refused.
        Calls that overlap on one device share that workspace and race."""
        self._check_buffers(d_img, None, d_out)
        _call("fddm_plan_execute_static", self.handle, d_img.ptr, d_out.ptr,
              stream.handle if stream is not None else None)

    def timing_begin(self, max_steps: int) -> None:
        _call("fddm_plan_timing_begin", self.handle, int(max_steps))

    def timing_end(self):
        ms = (ctypes.c_double * len(self.STAGES))()
        steps = ctypes.c_int(0)
        _call("fddm_plan_timing_end", self.handle, ms, ctypes.byref(steps))
        return dict(zip(self.STAGES, list(ms))), steps.value

    def spectra(self, d_img: DeviceBuffer, d_ws: DeviceBuffer) -> np.ndarray:
        L, H, W = self.shape
        nq = self.ny * (self.nx // 2 + 1)
        self._check_buffers(d_img, d_ws)
        out = np.empty((nq, L), dtype=np.complex128)
        _call("fddm_plan_spectra", self.handle, d_img.ptr, d_ws.ptr, out.ctypes.data)
        return out

    def destroy(self) -> None:
        if self.handle is not None and self.handle.value:
            lib().fddm_plan_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def synchronize() -> None:
    """Wait for the work queued on the current device's default stream."""
    _call("fddm_stream_synchronize", None)


def fill_uniform(buf: "DeviceBuffer", count: int, dtype, seed: int, lo: int, hi: int) -> None:
    """Fill ``count`` uint8/uint16 elements of a device buffer with uniform
    integers in [lo, hi) (device-side load generation)."""
    if count * np.dtype(dtype).itemsize > buf.nbytes:
        raise RuntimeError("fill_uniform: buffer too small")
    _call("fddm_fill_uniform", buf.ptr, count, DTYPE_CODES[np.dtype(dtype)], seed, lo, hi, None)
    _call("fddm_stream_synchronize", None)


class Server:
    """Continuous-batching request server on the current device.

    ``submit`` returns a request id at once; the request is admitted when a
    slot is free and its HBM footprint fits in the server's arena, and runs
    concurrently with the other admitted requests.  ``fetch`` copies a
    finished result to host, ``release`` frees it.
    """

    PENDING, RUNNING, DONE, FAILED = 0, 1, 2, 3

    def __init__(self, max_slots: int = 4, hbm_bytes: int = 0, aging_s: Optional[float] = None):
        h = ctypes.c_void_p()
        _call("fddm_server_create", ctypes.byref(h), max_slots, hbm_bytes)
        self._h = h
        self._keep = {}
        if aging_s is not None:
            lib().fddm_server_set_aging(self._h, aging_s)

    def submit(self, mode, img, lags, nx, ny, nt=0, window=None, precision="fp32c",
               out_dtype=np.float64, img_on_device=False, shape=None, dtype=None) -> int:
        """Queue one analysis.  ``img`` is a host array (read at admission; kept
        alive here until release) or, with ``img_on_device``, a DeviceBuffer
        holding a (L, H, W) stack of ``dtype`` given by ``shape``."""
        lg = _lags_arg(lags)
        if img_on_device:
            L, H, W = shape
            need = L * H * W * np.dtype(dtype).itemsize
            if img.nbytes < need:
                raise RuntimeError(f"input buffer holds {img.nbytes} bytes, a {shape} stack needs {need}")
            ptr, code = img.ptr, DTYPE_CODES[np.dtype(dtype)]
        else:
            img = _img_arg(img)
            L, H, W = img.shape
            ptr, code = img.ctypes.data, DTYPE_CODES[img.dtype]
        wp, wf32, wkeep = _window_arg(window, H, W)
        rid = ctypes.c_uint64()
        _call("fddm_server_submit", self._h, MODE_FFT if mode == "fft" else MODE_DIFF, ptr,
              1 if img_on_device else 0, code, L, H, W, lg.ctypes.data, len(lg), nx, ny, nt, wp,
              wf32, _prec_code(precision), _out_code(out_dtype), ctypes.byref(rid))
        self._keep[rid.value] = (img, (len(lg) + 2, ny, nx // 2 + 1), np.dtype(out_dtype))
        return rid.value

    def status(self, rid: int):
        st = ctypes.c_int()
        t = (ctypes.c_double * 3)()
        _call("fddm_server_status", self._h, rid, ctypes.byref(st), t)
        return st.value, (t[0], t[1], t[2])

    def wait(self, rid: int, timeout_s: float = -1.0) -> None:
        _call("fddm_server_wait", self._h, rid, timeout_s)

    def next_done(self, timeout_s: float = -1.0) -> int:
        """Id of the next finished request (completion order); 0 on timeout."""
        rid = ctypes.c_uint64()
        _call("fddm_server_next_done", self._h, timeout_s, ctypes.byref(rid))
        return rid.value

    def fetch(self, rid: int) -> np.ndarray:
        _, shape, dt = self._keep[rid]
        out = np.empty(shape, dtype=dt)
        _call("fddm_server_fetch", self._h, rid, out.ctypes.data)
        return out

    def release(self, rid: int) -> None:
        if lib().fddm_server_release(self._h, rid) != 0:
            raise RuntimeError(f"request {rid} cannot be released (unknown or still running)")
        self._keep.pop(rid, None)

    def stats(self) -> dict:
        v = (ctypes.c_uint64 * 7)()
        lib().fddm_server_stats(self._h, v)
        keys = ("pending", "running", "completed", "arena_bytes", "arena_used", "arena_peak", "plans")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fddm_server_destroy(self._h)
            self._h = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BatchItem(ctypes.Structure):
    """Mirror of ``fddm_batch_item`` (include/fddm_hip.h)."""

    _fields_ = [
        ("mode", ctypes.c_int), ("dtype", ctypes.c_int),
        ("length", ctypes.c_uint64), ("height", ctypes.c_uint64), ("width", ctypes.c_uint64),
        ("lags", ctypes.c_void_p), ("num_lags", ctypes.c_uint64),
        ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64), ("nt", ctypes.c_uint64),
        ("window", ctypes.c_void_p), ("window_f32", ctypes.c_int),
        ("precision", ctypes.c_int), ("out_dtype", ctypes.c_int),
        ("d_img", ctypes.c_void_p), ("d_out", ctypes.c_void_p),
    ]


class Batch:
    """Many device-resident stacks analysed by one graph launch.

    ``items``: dicts with keys mode, img (DeviceBuffer), shape (L, H, W),
    dtype, lags, nx, ny, nt (fft), window (optional), precision, out_dtype
    (optional) and out (DeviceBuffer of the plan's output size).
    """

    def __init__(self, items, autotune: bool = True):
        n = len(items)
        arr = (_BatchItem * n)()
        self._keep = []
        for k, it in enumerate(items):
            lg = _lags_arg(it["lags"])
            L, H, W = it["shape"]
            wp, wf32, wkeep = _window_arg(it.get("window"), H, W)
            self._keep.append((lg, wkeep, it["img"], it["out"]))
            need_in = L * H * W * np.dtype(it["dtype"]).itemsize
            need_out = (len(lg) + 2) * it["ny"] * (it["nx"] // 2 + 1) * np.dtype(
                it.get("out_dtype", np.float64)).itemsize
            if it["img"].nbytes < need_in or it["out"].nbytes < need_out:
                raise RuntimeError(f"batch item {k}: input or output buffer smaller than its sizes need "
                                   f"({it['img'].nbytes} < {need_in} or {it['out'].nbytes} < {need_out})")
            arr[k] = _BatchItem(
                MODE_FFT if it["mode"] == "fft" else MODE_DIFF, DTYPE_CODES[np.dtype(it["dtype"])],
                L, H, W, lg.ctypes.data, len(lg), it["nx"], it["ny"], it.get("nt", 0), wp, wf32,
                _prec_code(it.get("precision", "fp32c")), _out_code(it.get("out_dtype", np.float64)),
                it["img"].ptr, it["out"].ptr)
        self._n = n
        self._h = ctypes.c_void_p()
        _call("fddm_batch_create", ctypes.byref(self._h), ctypes.cast(arr, ctypes.c_void_p), n,
              1 if autotune else 0)

    def launch(self, stream: Optional["Stream"] = None) -> None:
        _call("fddm_batch_launch", self._h, None if stream is None else stream.handle)

    def launch_eager(self, stream: Optional["Stream"] = None) -> None:
        _call("fddm_batch_launch_eager", self._h, None if stream is None else stream.handle)

    def frames_per_block(self):
        v = (ctypes.c_int * self._n)()
        lib().fddm_batch_frames_per_block(self._h, v)
        return list(v)

    def spatial_kernels(self):
        """Per item, (rows, columns): True where the mixed-radix kernel runs."""
        r = (ctypes.c_int * self._n)()
        c = (ctypes.c_int * self._n)()
        lib().fddm_batch_spatial_kernels(self._h, r, c)
        return [(bool(a), bool(b)) for a, b in zip(r, c)]

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().fddm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
