"""Pins of the input-side helpers whose results the kernels consume, and the
``ddm()`` front end's error contract (CPU only).

* ``next_fast_len`` (reference src/python/_fftopt.py:29-150) against a
  brute-force enumeration of smooth numbers for every core;
* the windows (reference src/python/window.py:54-161) against their closed
  form: the DFT of a periodic cosine-sum window is non-zero only at bins
  0, +-1, .., +-(terms-1), with values N a_0, -N a_1 / 2, N a_2 / 2, ...;
* ``logspace_int`` against the examples in its reference docstring
  (src/python/lags.py:94-99), the only fixture the reference holds for it;
* ``_reconstruct_full_spectrum`` (reference imagestructurefunction.py:887-944)
  against fftshift(|fft2(x)|^2) of a real image, whose full plane is fixed by
  Hermitian symmetry;
* the ``ddm()`` error paths of tests/test_ddm.py:160-172 of the reference.
"""

import warnings

import numpy as np
import pytest
from scipy import fft as sfft

from fastddm_amd import _ddm
from fastddm_amd._fftopt import next_fast_len, next_power_of_2
from fastddm_amd.imagestructurefunction import ImageStructureFunction, _reconstruct_full_spectrum
from fastddm_amd.lags import fibonacci, logspace_int
from fastddm_amd.window import blackman, blackman_harris

LIMIT = 6000


def _smooth_table(limit, primes):
    ok = np.zeros(limit + 1, dtype=bool)
    for n in range(1, limit + 1):
        r = n
        for p in primes:
            while r % p == 0:
                r //= p
        ok[n] = r == 1
    return ok


SMOOTH7 = _smooth_table(2 * LIMIT, (2, 3, 5, 7))


def _brute(target, allowed):
    n = target
    while not allowed(n):
        n += 1
    return n


def _cpp_allowed(n):
    if SMOOTH7[n]:
        return True
    return (n % 11 == 0 and SMOOTH7[n // 11]) or (n % 13 == 0 and SMOOTH7[n // 13])


def test_next_power_of_2_closed_form():
    for t in range(1, 5000):
        p = next_power_of_2(t)
        assert p >= t and p & (p - 1) == 0 and (p == 1 or p // 2 < t)


@pytest.mark.parametrize("core", ["cuda", "hip", "CUDA", "Hip"])
def test_next_fast_len_7_smooth(core):
    for t in range(1, LIMIT):
        assert next_fast_len(t, core) == _brute(t, lambda n: SMOOTH7[n]), t


@pytest.mark.parametrize("core", ["cuda", "hip"])
def test_next_fast_len_force_even(core):
    # powers of two are returned as they are (1 included); otherwise the
    # smallest even 7-smooth number
    for t in range(1, LIMIT):
        expect = t if t & (t - 1) == 0 else _brute(t, lambda n: SMOOTH7[n] and n % 2 == 0)
        assert next_fast_len(t, core, force_even=True) == expect, t


def test_next_fast_len_cpp_11_13():
    for t in range(1, LIMIT):
        assert next_fast_len(t, "cpp") == _brute(t, _cpp_allowed), t
    # force_even has no effect on 'cpp' (reference :80-115)
    assert next_fast_len(11, "cpp", force_even=True) == 11


def test_next_fast_len_py_is_scipy():
    for t in range(1, 3000):
        p = next_power_of_2(t)
        assert next_fast_len(t, "py") == (p if p == t else sfft.next_fast_len(t))


def test_next_fast_len_config_sizes():
    # lengths the benchmark configurations use
    assert next_fast_len(2 * 2000, "hip") == 4000
    assert next_fast_len(2 * 4000, "hip") == 8000
    assert next_fast_len(2 * 300, "hip") == 600
    assert next_fast_len(2 * 64, "hip") == 128
    assert next_fast_len(200, "hip") == 200 and next_fast_len(150, "hip") == 150


def test_next_fast_len_unknown_core():
    with pytest.raises(ValueError):
        next_fast_len(100, "opencl")


# published coefficients (reference window.py:93 and :149)
BLACKMAN_A = (7938 / 18608, 9240 / 18608, 1430 / 18608)
HARRIS_A = (0.3635819, 0.4891775, 0.1365995, 0.0106411)


def _expected_dft(n, coeffs):
    spec = np.zeros(n, dtype=np.complex128)
    spec[0] = n * coeffs[0]
    for j in range(1, len(coeffs)):
        val = (-1) ** j * n * coeffs[j] / 2
        spec[j] += val
        spec[-j] += val
    return spec


@pytest.mark.parametrize("fn,coeffs", [(blackman, BLACKMAN_A), (blackman_harris, HARRIS_A)])
@pytest.mark.parametrize("shape", [(16, 16), (21, 33), (512, 512), (3, 7, 10)])
def test_window_closed_form(fn, coeffs, shape):
    w = fn(shape)
    ny, nx = shape[-2:]
    assert w.shape == (ny, nx) and w.dtype == np.float64
    # separable: each row is proportional to the first non-zero one
    wy, wx = w[:, 0] / np.sqrt(w[0, 0]), w[0, :] / np.sqrt(w[0, 0])
    assert np.allclose(w, np.multiply.outer(wy, wx), rtol=1e-12, atol=1e-15)
    for n, w1 in ((nx, wx), (ny, wy)):
        if n < 2 * len(coeffs):
            continue  # bins would alias
        assert np.allclose(np.fft.fft(w1), _expected_dft(n, coeffs), atol=1e-12 * n)
        # periodic: w(x) = w(N - x); w(0) = sum (-1)^j a_j
        assert np.allclose(w1[1:], w1[1:][::-1], atol=1e-15)
        assert np.isclose(w1[0], sum((-1) ** j * a for j, a in enumerate(coeffs)), atol=1e-15)
        if n % 2 == 0:  # centre value sum a_j (1 for both windows)
            assert np.isclose(w1[n // 2], sum(coeffs), atol=1e-15)


def test_window_centre_is_one():
    assert np.isclose(sum(BLACKMAN_A), 1.0, atol=1e-15)
    assert np.isclose(sum(HARRIS_A), 1.0, atol=1e-12)
    assert np.isclose(blackman_harris((64, 64))[32, 32], 1.0, atol=1e-12)


def test_logspace_int_reference_examples():
    # reference src/python/lags.py:94-99
    assert logspace_int(10, num=5, endpoint=True).tolist() == [1, 2, 3, 6, 10]
    assert logspace_int(10, num=5).tolist() == [1, 2, 3, 4, 7]


@pytest.mark.parametrize("stop,num", [(299, 50), (1999, 100), (50, 49), (4000, 80), (7, 6)])
def test_logspace_int_properties(stop, num):
    for endpoint in (True, False):
        v = logspace_int(stop, num=num, endpoint=endpoint)
        assert len(v) == num and v[0] == 1
        assert np.all(np.diff(v) >= 1), v  # distinct and increasing
        assert v[-1] <= stop
        if endpoint:
            assert v[-1] == stop
    # the diff-mode workload: 50 distinct lags in 1..299 ending at 299
    v = logspace_int(299, num=50, endpoint=True)
    assert len(np.unique(v)) == 50 and v.max() == 299


def test_fibonacci():
    assert fibonacci(10).tolist() == [1, 2, 3, 5, 8]
    assert fibonacci(8).tolist() == [1, 2, 3, 5]
    assert fibonacci(8, endpoint=True).tolist() == [1, 2, 3, 5, 8]


@pytest.mark.parametrize("ny,nx", [(8, 8), (9, 9), (8, 9), (9, 8), (6, 10), (7, 5)])
def test_full_spectrum_matches_fft2_of_real_image(ny, nx):
    rng = np.random.default_rng(ny * 100 + nx)
    x = rng.standard_normal((ny, nx))
    P = np.abs(np.fft.fft2(x)) ** 2
    half = np.fft.fftshift(P, axes=0)[:, : nx // 2 + 1]
    full = _reconstruct_full_spectrum(half, (ny, nx))
    assert np.allclose(full, np.fft.fftshift(P), rtol=1e-12, atol=1e-12)

    isf = ImageStructureFunction(np.stack([half, half, half]), 2 * np.pi * np.fft.fftfreq(nx)[: nx // 2 + 1],
                                 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny)), nx, ny, np.array([1.0]))
    assert np.allclose(isf.full_slice(0), np.fft.fftshift(P))
    assert np.allclose(isf.full_power_spec(), np.fft.fftshift(P))
    assert np.allclose(isf.full_kx(), 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(nx)))
    with pytest.raises(IndexError):
        isf.full_slice(1)


@pytest.fixture
def fake_hip(monkeypatch):
    """A registry entry that fails the test if the checks let a call through."""
    def never(*a, **k):
        raise AssertionError("backend called despite invalid arguments")

    monkeypatch.setitem(_ddm._backend, "hip", {"fft": never, "diff": never})
    return never


def test_ddm_negative_lags_raise(fake_hip):
    imgs = np.zeros((10, 8, 8), dtype=np.uint8)
    for mode in ("fft", "diff"):
        with pytest.raises(RuntimeError, match="Negative lags"):
            _ddm.ddm(imgs, [-1], core="hip", mode=mode)


def test_ddm_too_large_lags_raise(fake_hip):
    imgs = np.zeros((10, 8, 8), dtype=np.uint8)
    with pytest.raises(RuntimeError, match="larger than len"):
        _ddm.ddm(imgs, [1, 10], core="hip", mode="fft")


def test_ddm_window_shape_raises(fake_hip):
    imgs = np.zeros((10, 8, 6), dtype=np.uint16)
    with pytest.raises(RuntimeError, match="incompatible"):
        _ddm.ddm(imgs, [1, 2], core="hip", mode="fft", window=np.ones((9, 6)))


def test_ddm_unknown_core_and_mode(fake_hip):
    imgs = np.zeros((10, 8, 8), dtype=np.uint8)
    with pytest.raises(RuntimeError, match="Unknown core"):
        _ddm.ddm(imgs, [1], core="cuda")
    with pytest.raises(RuntimeError, match="Unknown mode"):
        _ddm.ddm(imgs, [1], core="hip", mode="wk")


def test_ddm_lag_zero_dropped_with_warning(monkeypatch):
    seen = {}

    def rec(img, lags, nx, ny, nt, window, **kw):
        seen["lags"] = lags
        seen["sizes"] = (nx, ny, nt)
        seen["window"] = window
        return np.zeros((len(lags) + 2, ny, nx // 2 + 1))

    monkeypatch.setitem(_ddm._backend, "hip", {"fft": rec, "diff": None})
    imgs = np.zeros((10, 9, 11), dtype=np.uint8)
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter("always")
        res = _ddm.ddm(imgs, [3, 0, 1, 3, 2], core="hip", mode="fft")
    assert any("Found 0 in lags" in str(x.message) for x in w)
    assert seen["lags"].tolist() == [1, 2, 3] and seen["lags"].dtype == np.int64
    assert seen["sizes"] == (12, 9, 20)
    assert seen["window"].size == 0
    assert res.tau.tolist() == [1.0, 2.0, 3.0]
    assert np.allclose(res.kx, 2 * np.pi * np.fft.fftfreq(12)[:7])
    assert np.allclose(res.ky, 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(9)))


def test_plain_fp32_warns(monkeypatch):
    from fastddm_amd import _core_hip, _ddm_hip

    monkeypatch.setattr(_core_hip, "set_device", lambda gpu_id: None)
    monkeypatch.setattr(_core_hip, "ddm_fft", lambda *a, **k: k["precision"])
    monkeypatch.setattr(_core_hip, "ddm_diff", lambda *a, **k: k["precision"])
    with pytest.warns(RuntimeWarning, match="fp32c"):
        assert _ddm_hip.ddm_fft_hip(None, [1], 8, 8, 16, None, precision="fp32") == "fp32"
    with pytest.warns(RuntimeWarning, match="fp32c"):
        assert _ddm_hip.ddm_diff_hip(None, [1], 8, 8, None, precision="fp32") == "fp32"
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert _ddm_hip.ddm_fft_hip(None, [1], 8, 8, 16, None, precision="fp32c") == "fp32c"
        assert _ddm_hip.ddm_fft_hip(None, [1], 8, 8, 16, None) == "fp64"


def test_single_precision_switch():
    """FASTDDM_SINGLE_PRECISION=1 selects float32 outputs (the reference's
    SINGLE_PRECISION build, _config.py.in:10-16), read at import."""
    import os
    import subprocess
    import sys

    code = "import numpy as np, fastddm_amd._config as c; print(c.DTYPE is np.float32, c.IS_SINGLE_PRECISION)"
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for val, expect in (("1", "True True"), ("0", "False False")):
        env = dict(os.environ, FASTDDM_SINGLE_PRECISION=val)
        out = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True,
                             check=True, timeout=120)
        assert out.stdout.strip() == expect


def test_wrapper_rejects_wrong_window_shape():
    from fastddm_amd import _core_hip

    with pytest.raises(RuntimeError, match="incompatible"):
        _core_hip._window_arg(np.ones((4, 5)), 5, 4)
    assert _core_hip._window_arg(None, 5, 4)[0] is None
    p, f32, keep = _core_hip._window_arg(np.ones((5, 4), dtype=np.float32), 5, 4)
    assert f32 == 1 and keep.shape == (5, 4)
