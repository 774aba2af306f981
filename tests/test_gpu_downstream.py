"""The consumers of a device-produced result, and the single-precision build
through the ddm() front end.

* Camera-noise estimators (every mode, on the image structure function and
  on its azimuthal average) and the ISF estimate run on the HIP result of
  the reference's confocal fixture and on the FP64 oracle's result wrapped
  the same way; they agree (fp64 to 1e-9, fp32c to the single-precision
  bar 1e-3; the polyfit intercept to 1e-7 resp. 1e-2 of |D(q, 1)|).
* FASTDDM_SINGLE_PRECISION=1 (the reference's SINGLE_PRECISION build,
  _config.py.in:10-16) with single-precision arithmetic (fp32c): float32
  results whose
  every element is within rtol 1e-3, atol 0 of the FP64 oracle, as the
  reference's single-precision cross-core tests require
  (tests/test_ddm.py:116-157).  Run in a child process, since DTYPE is read
  when the package is imported.
"""

import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu

LAGS = np.arange(1, 10)
MODES = ("zero", "min", "high_q", "power_spec", "var", "polyfit")


def _check_noise(got, ref, d1, mode, rtol, ptol):
    for g, r in zip(got, ref):
        g, r = np.asarray(g, dtype=np.float64), np.asarray(r, dtype=np.float64)
        if mode == "polyfit":
            bound = ptol * (np.abs(d1) + np.abs(r))
        else:
            bound = rtol * np.abs(r)
        ok = (np.abs(g - r) <= bound) | (np.isnan(g) & np.isnan(r))
        assert ok.all(), (mode, np.nanmax(np.abs(g - r) / (np.abs(r) + 1e-300)))


def _oracle_isf(confocal, mode):
    from fastddm_amd import ImageStructureFunction

    L, H, W = confocal.shape
    if mode == "fft":
        data = onp.ddm_fft(confocal, LAGS, W, H, 2 * L, subtract_mean=True)
    else:
        data = onp.ddm_diff(confocal, LAGS, W, H)
    kx = 2 * np.pi * np.fft.fftfreq(W)[: W // 2 + 1]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(H))
    return ImageStructureFunction(data, kx, ky, W, H, LAGS.astype(np.float64))


@pytest.mark.parametrize("precision,rtol", [("fp64", 1e-9), ("fp32c", 1e-3)])
def test_noise_and_isf_estimates_on_device_result(hip, confocal, precision, rtol):
    import fastddm_amd as fddm
    from fastddm_amd.intermediatescatteringfunction import azavg2isf_estimate
    from fastddm_amd.noise_est import estimate_camera_noise

    got = fddm.ddm(confocal, LAGS, core="hip", mode="fft", precision=precision)
    ref = _oracle_isf(confocal, "fft")
    assert np.isclose(got._data, ref._data, rtol=rtol, atol=0).all()
    # the polyfit intercept extrapolates the first lags to zero lag and
    # amplifies their errors (about 30x on this stack): it is held to a
    # fraction of |D(q, 1)| instead of its own (possibly near-zero) value
    ptol = 1e-2 if precision == "fp32c" else 1e-7
    for mode in MODES:
        _check_noise(estimate_camera_noise(got, mode), estimate_camera_noise(ref, mode), ref.data[0], mode, rtol, ptol)
    aa_g = fddm.azimuthal_average(got, bins=20)
    aa_r = fddm.azimuthal_average(ref, bins=20)
    assert np.allclose(aa_g.data, aa_r.data, rtol=rtol, equal_nan=True)
    for mode in MODES:
        _check_noise(estimate_camera_noise(aa_g, mode), estimate_camera_noise(aa_r, mode), aa_r.data[:, 0], mode,
                     rtol, ptol)
    isf_g = azavg2isf_estimate(aa_g)
    isf_r = azavg2isf_estimate(aa_r)
    assert np.allclose(isf_g.data, isf_r.data, rtol=10 * ptol, atol=10 * ptol, equal_nan=True)


_CHILD = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import fastddm_amd as fddm
from fastddm_amd._config import DTYPE
assert DTYPE is np.float32, DTYPE
with np.load(sys.argv[2]) as z:
    frames = z["frames"]
lags = np.arange(1, 10)
out = {}
for mode in ("fft", "diff"):
    r = fddm.ddm(frames, lags, core="hip", mode=mode, precision="fp32c")
    assert r._data.dtype == np.float32 and r.kx.dtype == np.float32 and r.tau.dtype == np.float32
    out[mode] = r._data
np.savez(sys.argv[3], **out)
print("child ok", flush=True)
"""


def test_single_precision_front_end(hip, confocal, tmp_path):
    dst = str(tmp_path / "single.npz")
    env = dict(os.environ, FASTDDM_SINGLE_PRECISION="1")
    res = subprocess.run([sys.executable, "-c", _CHILD, ROOT, os.path.join(GOLDEN, "confocal_20.npz"), dst],
                         env=env, capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stderr[-2000:]
    with np.load(dst) as z:
        for mode in ("fft", "diff"):
            ref = _oracle_isf(confocal, mode)._data
            got = z[mode]
            assert got.dtype == np.float32
            assert np.isclose(got, ref, rtol=1e-3, atol=0).all(), (mode, onp.relative_error(got, ref).max())
