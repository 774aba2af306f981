"""The C-ABI library loads and exports every entry point include/fddm_hip.h
declares; argument errors come back as codes + messages (no GPU needed, no
compute calls)."""

import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "fddm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(fddm_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from fastddm_amd import _build, _core_hip

    if not os.path.isfile(_build.TARGET):
        _build.build(verbose=False)
    return _core_hip.lib()


def test_header_declares_the_expected_entry_points():
    names = declared_functions()
    for must in ("fddm_ddm_fft", "fddm_ddm_diff", "fddm_set_device", "fddm_get_num_devices",
                 "fddm_get_device", "fddm_get_free_device_memory", "fddm_plan_create",
                 "fddm_plan_execute"):
        assert must in names


def test_library_exports_every_declared_symbol(lib):
    from fastddm_amd import _core_hip

    names = declared_functions()
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(_core_hip.EXPORTED_SYMBOLS) == names


def _plan_create(lib, **over):
    args = dict(mode=1, dtype=2, L=8, H=6, W=6, lags=np.array([1, 2], np.uint32), nx=6, ny=6,
                nt=16, precision=0, out=0)
    args.update(over)
    plan = ctypes.c_void_p()
    err = ctypes.create_string_buffer(512)
    lg = args["lags"]
    rc = lib.fddm_plan_create(ctypes.byref(plan), args["mode"], args["dtype"], args["L"], args["H"],
                              args["W"], lg.ctypes.data, len(lg), args["nx"], args["ny"], args["nt"],
                              None, 0, args["precision"], args["out"], err, 512)
    return rc, err.value.decode()


@pytest.mark.parametrize("over,code,text", [
    (dict(mode=7), 1, "unknown mode"),
    (dict(dtype=42), 1, "unknown input dtype"),
    (dict(precision=9), 1, "unknown precision"),
    (dict(out=3), 1, "unknown output dtype"),
    (dict(lags=np.array([0, 1], np.uint32)), 1, "outside"),
    (dict(lags=np.array([8], np.uint32)), 1, "outside"),
    (dict(lags=np.array([2, 1], np.uint32)), 1, "sorted"),
    (dict(nx=4), 1, "nx, ny"),
    (dict(nx=11, W=11), 2, "2^a 3^b 5^c 7^d"),
    (dict(nt=10), 1, "nt must be"),
    # beyond the four-step time transform: 2^17 points (fp32), 2^16 (fp64)
    (dict(L=70000, nt=140000, precision=1), 2, "too long"),
    (dict(L=40000, nt=80000, precision=0), 2, "too long"),
])
def test_plan_argument_errors_are_reported(lib, over, code, text):
    rc, msg = _plan_create(lib, **over)
    assert rc == code
    assert text in msg


def test_host_memory_utilities(lib):
    """get_free_host_memory / get_host_memory_* / check_host_memory_* (the
    reference's device utilities, src/cuda/memchk_gpu.cc:35-199) are host-only
    and answer without a GPU."""
    from fastddm_amd import _core_hip as hip

    with open("/proc/meminfo") as f:
        avail = next(int(l.split()[1]) * 1024 for l in f if l.startswith("MemAvailable:"))
    got = hip.get_free_host_memory()
    assert abs(got - avail) < 0.25 * avail  # the two reads are moments apart
    stage = 2 * (64 << 20)
    assert hip.get_host_memory_fft(512, 512, 2000, 1999) == 257 * 512 * 2001 * 8 + stage
    assert hip.get_host_memory_diff(512, 512, 300, 50, np.float32) == 257 * 512 * 52 * 4 + stage
    assert hip.check_host_memory_fft(64, 64, 100, 99)
    assert not hip.check_host_memory_diff(1 << 20, 1 << 20, 100, 1 << 20)
    err = ctypes.create_string_buffer(256)
    b = ctypes.c_uint64(0)
    assert lib.fddm_get_host_memory(7, 8, 8, 4, 3, 0, ctypes.byref(b), err, 256) != 0
    assert b"unknown mode" in err.value
