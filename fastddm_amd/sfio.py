"""``.sf.ddm`` files: the binary format of an image structure function.

Byte-compatible with the reference writer and reader, file version 0.3
(``SFWriter`` / ``SFReader`` / ``SFParser``, src/python/imagestructurefunction.py:409-884;
``Writer`` / ``Reader`` / ``Parser``, src/python/_io_common.py):

* header, 64 bytes: ``"LL"``/``"BB"`` byte order, file id 73 (``H``),
  version (major, minor) (``BB``), dtype ``"d"``/``"f"``, then ``Q`` values
  Nt, Ny, Nx, Nextra, full width, full height, zero padding;
* from byte 64: ``_data`` (Nt + Nextra slices of Ny x Nx, C order), then kx
  (Nx), ky (Ny), tau (Nt), pixel_size, delta_t, all in the data dtype;
* trailer, ``Q`` byte offsets written as extra, delta_t, pixel_size, tau, ky,
  kx, data -- so that the data offset is the last 8 bytes of the file.

``save_device_result`` streams a result that is still in device memory
(``(NL + 2, ny, nx//2 + 1)``, e.g. a plan's output or a served request's
result) to the file through a bounded host buffer, without a full host copy.
"""

import ctypes
import os
import struct
import sys
from typing import Optional, Tuple

import numpy as np

from .imagestructurefunction import ImageStructureFunction, _reconstruct_full_spectrum

VERSION = (0, 3)
HEAD_BYTE_LEN = 64
FILE_ID = 73
_FMT = {np.dtype(np.float64): "d", np.dtype(np.float32): "f"}


def _fmt_of(dtype) -> str:
    dt = np.dtype(dtype)
    if dt not in _FMT:
        raise NotImplementedError(f"Numpy dtype to binary format converter not implemented for type {dt.name}.")
    return _FMT[dt]


def _header(nt, ny, nx, nextra, width, height, fmt) -> bytes:
    h = ("LL" if sys.byteorder == "little" else "BB").encode("utf-8")
    h += struct.pack("H", FILE_ID) + struct.pack("BB", *VERSION) + fmt.encode("utf-8")
    h += struct.pack("QQQQQQ", nt, ny, nx, nextra, width, height)
    return h + bytes(HEAD_BYTE_LEN - len(h))


def _tail(fh, nt, ny, nx, nextra, fmt, kx, ky, tau, pixel_size, delta_t) -> None:
    """kx, ky, tau, pixel size, frame interval and the offset trailer."""
    sz = struct.calcsize(fmt)
    dt = np.dtype(fmt)
    np.asarray(kx, dtype=dt).tofile(fh)
    np.asarray(ky, dtype=dt).tofile(fh)
    np.asarray(tau, dtype=dt).tofile(fh)
    fh.write(struct.pack(fmt, pixel_size))
    fh.write(struct.pack(fmt, delta_t))
    data_offset = HEAD_BYTE_LEN
    extra_offset = data_offset + nt * ny * nx * sz
    kx_offset = extra_offset + nextra * ny * nx * sz
    ky_offset = kx_offset + nx * sz
    tau_offset = ky_offset + ny * sz
    pixel_size_offset = tau_offset + nt * sz
    delta_t_offset = pixel_size_offset + sz
    for off in (extra_offset, delta_t_offset, pixel_size_offset, tau_offset, ky_offset, kx_offset, data_offset):
        fh.write(struct.pack("Q", off))


def _file_name(fname: str) -> str:
    d, name = os.path.split(fname)
    name = name if name.endswith(".sf.ddm") else f"{name}.sf.ddm"
    if d:
        os.makedirs(d, exist_ok=True)
    return os.path.join(d, name)


def save(isf: ImageStructureFunction, fname: str = "analysis_blob") -> str:
    """Write ``isf`` to ``fname`` (``.sf.ddm`` appended when missing), as
    ``ImageStructureFunction.save`` of the reference (:279-293).  Returns the
    path written."""
    fmt = _fmt_of(isf._data.dtype)
    nt, ny, nx = isf.shape
    nextra = len(isf._data) - nt
    path = _file_name(fname)
    with open(path, "wb") as fh:
        fh.write(_header(nt, ny, nx, nextra, isf.width, isf.height, fmt))
        np.ascontiguousarray(isf._data).tofile(fh)
        _tail(fh, nt, ny, nx, nextra, fmt, isf.kx, isf.ky, isf.tau, isf.pixel_size, isf.delta_t)
    return path


def save_device_result(fname: str, d_data, shape: Tuple[int, int, int], kx, ky, tau, width: int, height: int,
                       dtype=np.float64, pixel_size: float = 1.0, delta_t: float = 1.0,
                       chunk_bytes: int = 64 << 20) -> str:
    """Stream a device-resident ``(NL + 2, ny, nxh)`` result to ``fname`` in
    chunks of at most ``chunk_bytes`` (device -> host -> file)."""
    from . import _core_hip as core

    fmt = _fmt_of(dtype)
    rows, ny, nx = shape
    nt = len(tau)
    if rows != nt + 2:
        raise ValueError("shape[0] must be len(tau) + 2")
    total = rows * ny * nx * np.dtype(dtype).itemsize
    base = d_data.ptr.value if hasattr(d_data, "ptr") else int(d_data)
    host = np.empty(min(total, max(int(chunk_bytes), 4096)), dtype=np.uint8)
    path = _file_name(fname)
    with open(path, "wb") as fh:
        fh.write(_header(nt, ny, nx, rows - nt, width, height, fmt))
        off = 0
        while off < total:
            m = min(len(host), total - off)
            core._call("fddm_memcpy_dtoh", host.ctypes.data, ctypes.c_void_p(base + off), m)
            fh.write(memoryview(host)[:m])
            off += m
        _tail(fh, nt, ny, nx, rows - nt, fmt, kx, ky, tau, pixel_size, delta_t)
    return path


class SFReader:
    """Reader of ``.sf.ddm`` files (reference ``SFReader`` / ``SFParser``)."""

    def __init__(self, file: str):
        self._fh = open(file, "rb")
        head = self._fh.read(HEAD_BYTE_LEN)
        flag = head[:2].decode("utf-8")
        if flag not in ("LL", "BB"):
            self._fh.close()
            raise RuntimeError(f"Byteorder {flag} not recognized. Input file might be corrupted.")
        self._bo = "<" if flag == "LL" else ">"
        file_id = struct.unpack(self._bo + "H", head[2:4])[0]
        if file_id != FILE_ID:
            self._fh.close()
            raise RuntimeError(f"File identifier {file_id} not compatible with image structure function file "
                               f"({FILE_ID}). Input file might be wrong or corrupted.")
        self.version = tuple(head[4:6])
        self.dtype = head[6:7].decode("utf-8")
        nt, ny, nx, nextra, width, height = struct.unpack(self._bo + "QQQQQQ", head[7:55])
        self._fh.seek(-7 * 8, os.SEEK_END)
        (extra, delta_t_off, pixel_size_off, tau_off, ky_off, kx_off,
         data_off) = struct.unpack(self._bo + "QQQQQQQ", self._fh.read(56))
        self._meta = dict(Nt=nt, Ny=ny, Nx=nx, Nextra=nextra, width=width, height=height, data_offset=data_off,
                          extra_offset=extra, kx_offset=kx_off, ky_offset=ky_off, tau_offset=tau_off)
        self._meta["pixel_size"] = self._value(pixel_size_off)
        self._meta["delta_t"] = self._value(delta_t_off)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def close(self) -> None:
        if self._fh:
            self._fh.close()
            self._fh = None

    @property
    def metadata(self) -> dict:
        return dict(self._meta)

    def _value(self, offset: int) -> float:
        self._fh.seek(offset)
        fmt = self._bo + self.dtype
        return struct.unpack(fmt, self._fh.read(struct.calcsize(fmt)))[0]

    def _array(self, offset: int, shape) -> np.ndarray:
        self._fh.seek(0)
        count = int(np.prod(shape))
        return np.fromfile(self._fh, self._bo + self.dtype, count, offset=offset).reshape(shape)

    def load(self) -> ImageStructureFunction:
        if self.version != VERSION:
            raise IOError(f"File version {self.version} not supported.")
        m = self._meta
        data = self._array(m["data_offset"], (m["Nt"] + m["Nextra"], m["Ny"], m["Nx"]))
        return ImageStructureFunction(data, self.get_kx(False), self.get_ky(), m["width"], m["height"],
                                      self.get_tau(), m["pixel_size"], m["delta_t"])

    def get_kx(self, full: Optional[bool] = True) -> np.ndarray:
        m = self._meta
        kx = self._array(m["kx_offset"], m["Nx"])
        if not full:
            return kx
        out = np.zeros_like(kx, shape=m["width"])
        out[: m["Nx"]] = kx
        out[m["Nx"]:] = -np.flip(kx[1:-1]) if m["width"] % 2 == 0 else -np.flip(kx[1:])
        return np.fft.fftshift(out)

    def get_ky(self, full: Optional[bool] = True) -> np.ndarray:
        return self._array(self._meta["ky_offset"], self._meta["Ny"])

    def get_tau(self) -> np.ndarray:
        return self._array(self._meta["tau_offset"], self._meta["Nt"])

    def _slice(self, offset: int, full: bool) -> np.ndarray:
        m = self._meta
        half = self._array(offset, (m["Ny"], m["Nx"]))
        return _reconstruct_full_spectrum(half, (m["height"], m["width"])) if full else half

    def get_frame(self, index: int, full: Optional[bool] = True) -> np.ndarray:
        m = self._meta
        if index < 0 or index >= m["Nt"]:
            raise IndexError(f"Index out of range. Choose an index between 0 and {m['Nt'] - 1}.")
        sz = struct.calcsize(self.dtype)
        return self._slice(m["data_offset"] + index * m["Nx"] * m["Ny"] * sz, full)

    def get_power_spec(self, full: Optional[bool] = True) -> np.ndarray:
        return self._slice(self._meta["extra_offset"], full)

    def get_var(self, full: Optional[bool] = True) -> np.ndarray:
        m = self._meta
        return self._slice(m["extra_offset"] + m["Nx"] * m["Ny"] * struct.calcsize(self.dtype), full)


def load(fname: str) -> ImageStructureFunction:
    """Read an ``ImageStructureFunction`` from a ``.sf.ddm`` file."""
    with SFReader(fname) as r:
        return r.load()
