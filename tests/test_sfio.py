"""``.sf.ddm`` writer/reader: the byte layout of the reference's file version
0.3 (header fields at fixed offsets, data, axes, trailer of offsets read from
the end) checked field by field, round trips, and the full-plane readers."""

import os
import struct
import sys

import numpy as np
import pytest

from fastddm_amd import ImageStructureFunction
from fastddm_amd import sfio


def _isf(dtype=np.float64, nl=5, ny=6, nxh=4, width=6, seed=0):
    rng = np.random.default_rng(seed)
    data = rng.random((nl + 2, ny, nxh)).astype(dtype)
    kx = (2 * np.pi * np.fft.fftfreq(width)[:nxh]).astype(dtype)
    ky = (2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny))).astype(dtype)
    tau = np.arange(1, nl + 1).astype(dtype)
    return ImageStructureFunction(data, kx, ky, width, ny, tau, 0.5, 0.25)


@pytest.mark.parametrize("dtype,fmt", [(np.float64, "d"), (np.float32, "f")])
def test_byte_layout(tmp_path, dtype, fmt):
    isf = _isf(dtype)
    path = isf.save(str(tmp_path / "blob"))
    assert path.endswith("blob.sf.ddm")
    raw = open(path, "rb").read()
    sz = np.dtype(dtype).itemsize
    nt, ny, nx, nextra = 5, 6, 4, 2
    assert raw[:2] == (b"LL" if sys.byteorder == "little" else b"BB")
    bo = "<" if sys.byteorder == "little" else ">"
    assert struct.unpack(bo + "H", raw[2:4])[0] == 73
    assert tuple(raw[4:6]) == (0, 3)
    assert raw[6:7] == fmt.encode()
    assert struct.unpack(bo + "QQQQQQ", raw[7:55]) == (nt, ny, nx, nextra, 6, 6)
    assert raw[55:64] == bytes(9)
    data_bytes = (nt + nextra) * ny * nx * sz
    assert np.array_equal(np.frombuffer(raw[64:64 + data_bytes], dtype).reshape(isf._data.shape), isf._data)
    off_kx = 64 + data_bytes
    assert np.array_equal(np.frombuffer(raw[off_kx:off_kx + nx * sz], dtype), isf.kx)
    off_tau = off_kx + (nx + ny) * sz
    assert np.array_equal(np.frombuffer(raw[off_tau:off_tau + nt * sz], dtype), isf.tau)
    ps, dt = struct.unpack(bo + fmt * 2, raw[off_tau + nt * sz:off_tau + (nt + 2) * sz])
    assert (ps, dt) == (0.5, 0.25)
    trailer = struct.unpack(bo + "7Q", raw[-56:])
    extra = 64 + nt * ny * nx * sz
    expect = (extra, off_tau + (nt + 1) * sz, off_tau + nt * sz, off_tau, off_kx + nx * sz, off_kx, 64)
    assert trailer == expect
    assert len(raw) == off_tau + (nt + 2) * sz + 56
    assert extra == 64 + nt * ny * nx * sz


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("width", [6, 7])
def test_round_trip_and_full_plane(tmp_path, dtype, width):
    isf = _isf(dtype, nxh=width // 2 + 1, width=width, seed=width)
    path = isf.save(str(tmp_path / "sub" / "x.sf.ddm"))
    back = sfio.load(path)
    assert np.array_equal(back._data, isf._data) and back._data.dtype == isf._data.dtype
    for attr in ("kx", "ky", "tau"):
        assert np.array_equal(getattr(back, attr), getattr(isf, attr))
    assert (back.width, back.height, back.pixel_size, back.delta_t) == (width, 6, 0.5, 0.25)
    with sfio.SFReader(path) as r:
        assert np.allclose(r.get_kx(), np.fft.fftshift(2 * np.pi * np.fft.fftfreq(width)))
        assert np.array_equal(r.get_frame(2), isf.full_slice(2))
        assert np.array_equal(r.get_frame(2, full=False), isf.data[2])
        assert np.array_equal(r.get_power_spec(), isf.full_power_spec())
        assert np.array_equal(r.get_var(), isf.full_var())
        with pytest.raises(IndexError):
            r.get_frame(5)


def test_rejects_foreign_files(tmp_path):
    p = tmp_path / "bad.sf.ddm"
    p.write_bytes(b"XX" + bytes(200))
    with pytest.raises(RuntimeError, match="Byteorder"):
        sfio.SFReader(str(p))
    raw = bytearray(open(_isf().save(str(tmp_path / "ok")), "rb").read())
    raw[2:4] = struct.pack("<H", 74)
    p.write_bytes(bytes(raw))
    with pytest.raises(RuntimeError, match="identifier"):
        sfio.SFReader(str(p))
    with pytest.raises(NotImplementedError):
        _isf(np.float16).save(str(tmp_path / "h"))
