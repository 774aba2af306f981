"""The benchmark's workload at its full size -- 2000 frames of 512 x 512
uint16, lags 1..1999, Blackman-Harris window, nt = 4000 -- checked through
properties that hold at any size, since the oracle cannot run there in
seconds:

* the compensated FP32 path agrees with the FP64 path to a median relative
  error of 1e-5, and every element within the reference's single-precision
  tolerance, rtol 1e-3 with atol 0 (tests/test_ddm.py:116-157 of the
  reference);
* D is a mean of squared moduli: positive for these noisy frames, and
  PS >= VAR >= 0;
* direct differences (diff mode) on a subset of lags reproduce the
  Wiener-Khinchin (fft mode) values in FP64.
"""

import numpy as np
import pytest

from fastddm_amd import window as fwin

pytestmark = pytest.mark.gpu

L, H, W, NT = 2000, 512, 512, 4000


def _run(hip, d_img, mode, lags, precision):
    win = fwin.blackman_harris((H, W))
    plan = hip.Plan(mode, np.uint16, (L, H, W), lags, W, H, NT if mode == "fft" else 0, win, precision)
    d_out = hip.DeviceBuffer(plan.output_size())
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    plan.execute(d_img, d_out, d_ws)
    out = d_out.to_array(plan.out_shape, np.float64)
    plan.destroy()
    d_out.free()
    d_ws.free()
    return out


def test_full_size_properties(hip):
    print(f"full size: {L} x {H} x {W} uint16, all lags", flush=True)
    d_img = hip.DeviceBuffer(L * H * W * 2)
    hip.fill_uniform(d_img, L * H * W, np.uint16, 11, 800, 1200)
    lags = np.arange(1, L)
    ref = _run(hip, d_img, "fft", lags, "fp64")
    got = _run(hip, d_img, "fft", lags, "fp32c")
    assert ref.shape == got.shape == (L + 1, H, W // 2 + 1)

    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    assert np.median(rel) <= 1e-5, np.median(rel)
    assert rel.max() <= 1e-3, rel.max()
    assert np.array_equal(got == 0, ref == 0)
    del got, rel, nz

    D, PS, VAR = ref[:-2], ref[-2], ref[-1]
    assert np.all(D > 0)
    assert np.all(VAR >= 0) and np.all(PS >= VAR)

    sub = np.array([1, 2, 17, 640, 1500, 1999])
    direct = _run(hip, d_img, "diff", sub, "fp64")
    assert np.allclose(direct[:-2], D[sub - 1], rtol=1e-9, atol=0)
    assert np.allclose(direct[-2:], ref[-2:], rtol=1e-9, atol=0)
    d_img.free()


def test_large_size_properties(hip):
    """4000 frames of 1024 x 1024 uint16 (nt = 8000: the two-column N = 8192
    time plan, 8-frame blocks whose 64-byte runs take the XCD-aware order,
    FP64 rows on 4-frame blocks), about 80 log-spaced lags from 1 to 3999 and
    the last 80: the same properties as above."""
    Lb, Hb, Wb = 4000, 1024, 1024
    print(f"large size: {Lb} x {Hb} x {Wb} uint16, log-spaced lags", flush=True)
    d_img = hip.DeviceBuffer(Lb * Hb * Wb * 2)
    hip.fill_uniform(d_img, Lb * Hb * Wb, np.uint16, 21, 800, 1200)
    sub = np.array([1, 37, 1000, Lb - 1])  # also checked by direct differences
    # log-spaced, plus every lag with 1..80 frame pairs: the few-pair lags are
    # where the float transforms' rounding is largest relative to D
    lags = np.unique(np.concatenate([np.geomspace(1, Lb - 1, 100).astype(int), np.arange(Lb - 80, Lb), sub]))
    win = fwin.blackman_harris((Hb, Wb))

    def run(mode, lg, precision):
        plan = hip.Plan(mode, np.uint16, (Lb, Hb, Wb), lg, Wb, Hb, 2 * Lb if mode == "fft" else 0, win, precision)
        d_out = hip.DeviceBuffer(plan.output_size())
        d_ws = hip.DeviceBuffer(plan.workspace_size())
        plan.execute(d_img, d_out, d_ws)
        out = d_out.to_array(plan.out_shape, np.float64)
        plan.destroy()
        d_out.free()
        d_ws.free()
        return out

    ref = run("fft", lags, "fp64")
    got = run("fft", lags, "fp32c")
    assert ref.shape == got.shape == (len(lags) + 2, Hb, Wb // 2 + 1)
    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    assert np.median(rel) <= 1e-5, np.median(rel)
    assert rel.max() <= 1e-3, rel.max()
    assert np.array_equal(got == 0, ref == 0)
    del got, rel, nz

    D, PS, VAR = ref[:-2], ref[-2], ref[-1]
    assert np.all(D > 0)
    assert np.all(VAR >= 0) and np.all(PS >= VAR)

    direct = run("diff", sub, "fp64")
    idx = np.searchsorted(lags, sub)
    assert np.all(lags[idx] == sub)
    assert np.allclose(direct[:-2], D[idx], rtol=1e-9, atol=0)
    assert np.allclose(direct[-2:], ref[-2:], rtol=1e-9, atol=0)
    d_img.free()


def test_benchmark_speckle_precision(hip):
    """The benchmark's own input (2000 x 512^2 uint16 AR(1) speckle, sigma 1.5,
    D 0.2, mean 1000, std 150, seed 2), all lags, Blackman-Harris: slow low-q
    modes whose D at large lags is small against their variance -- where the
    Wiener-Khinchin rounding of the float transforms is largest relative to
    D -- and chance near-recurrences at the last lag.  fp32c within rtol 1e-3
    of fp64 everywhere, median 1e-5."""
    from fastddm_amd.speckle import CONFIGS, ar1_speckle

    img = ar1_speckle(**CONFIGS["precision_512"])
    print(f"benchmark speckle: {img.shape} {img.dtype}", flush=True)
    d_img = hip.DeviceBuffer.from_array(img)
    del img
    lags = np.arange(1, L)
    ref = _run(hip, d_img, "fft", lags, "fp64")
    got = _run(hip, d_img, "fft", lags, "fp32c")
    d_img.free()
    nz = ref != 0
    rel = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    print(f"benchmark speckle: fp32c vs fp64 median {np.median(rel):.2e} max {rel.max():.2e}", flush=True)
    assert np.median(rel) <= 1e-5, np.median(rel)
    assert rel.max() <= 1e-3, rel.max()
