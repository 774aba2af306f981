"""Compensated FP32 (fp32c) at the reference's single-precision bar on inputs
built so that the float rounding would otherwise break it.

fp32c stores float spectra and runs the time transforms in float; D is then
accurate to about 1e-7 of the spectrum (spatial transform) and of
sum_t |G|^2 (Wiener-Khinchin), which is far below 1e-3 of D except where D
is tiny against those.  Two corrections cover that:

* lags with at most 8 frame pairs are recomputed from FP64 spectra of the
  first and last 8 frames -- here the last frame repeats the first one plus a
  strong cosine and weak noise, so at every q off the cosine the single
  difference is 1e-7 of the frame's spectrum;
* lags whose Wiener-Khinchin value is small against sum_t |G|^2 are
  recomputed by direct differences of the float series -- here the stack is
  periodic in time up to weak noise, so D at multiples of the period is
  1e-5 of the variance.

Both are checked against the FP64 oracle's direct differences with
np.isclose(rtol=1e-3, atol=0) on every element (the reference's
single-precision bar, tests/test_ddm.py:116-157).
"""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu


def _check(out, ref):
    rel = onp.relative_error(out, ref)
    assert np.isclose(out, ref, rtol=1e-3, atol=0).all(), rel.max()
    return rel


def test_last_lag_near_recurrence_fp64_edges(hip):
    rng = np.random.default_rng(41)
    L, H, W = 40, 64, 64
    img = rng.normal(20000.0, 30.0, (L, H, W))
    x = np.arange(W)
    img[-1] = img[0] + 10000.0 * np.cos(2 * np.pi * 5 * x / W)[None, :] + rng.normal(0.0, 0.5, (H, W))
    img = np.rint(img).astype(np.uint16)
    lags = np.arange(1, L)
    ref = onp.ddm_diff(img, lags, W, H)
    print("near recurrence at the last lag: fft and diff, fp32c", flush=True)
    out = hip.ddm_fft(img, lags, W, H, next_fast_len(2 * L, "hip"), None, precision="fp32c")
    rel = _check(out, ref)
    assert np.median(rel) <= 1e-5
    _check(hip.ddm_diff(img, lags, W, H, None, precision="fp32c"), ref)
    # the last lag really is 1e-7 of the frame's spectrum here
    F = onp.spectra(img.astype(np.float64) - img[0].astype(np.float64), W, H)
    assert np.median(np.abs(F[-1]) ** 2) < 1e-6 * np.mean(np.abs(F[-1]) ** 2)


@pytest.mark.parametrize("L,H,W", [(400, 48, 40), (1200, 32, 32), (2200, 16, 16)])
def test_periodic_stack_recomputed_lags(hip, L, H, W):
    rng = np.random.default_rng(L + H)
    P = 100
    base = rng.normal(10000.0, 1000.0, (P, H, W))
    img = np.rint(np.tile(base, (L // P, 1, 1)) + rng.normal(0.0, 3.0, (L, H, W))).astype(np.uint16)
    lags = np.arange(1, L)
    ref = onp.ddm_diff(img, lags, W, H)
    print(f"periodic stack {L} x {H} x {W}, period {P}: fp32c", flush=True)
    out = hip.ddm_fft(img, lags, W, H, next_fast_len(2 * L, "hip"), None, precision="fp32c")
    rel = _check(out, ref)
    assert np.median(rel) <= 1e-5
    # D at multiples of the period is the weak noise only
    per = ref[P - 1 :: P][: L // P - 1, :, :]
    assert np.median(per / ref[-1]) < 1e-4
