"""Device azimuthal average against the FP64 oracle (same bin maps), on real
structure functions: int and explicit bins, range, mask, weights, float32
data, no errors, and a result that never leaves the device."""

import numpy as np
import pytest

import fastddm_amd as fddm
from fastddm_amd.azimuthalaverage import _half_plane_counts, _k_modulus, azimuthal_average_device, bin_maps
from fastddm_amd.speckle import ar1_speckle
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu


def _ref(isf, bins, rng=None, mask=None, weights=None, eval_err=True):
    dist = _k_modulus(isf.kx, isf.ky)
    counts = _half_plane_counts(dist.shape, isf.width)
    bi, w, w2, edges, k = bin_maps(dist, bins, rng, mask, weights, counts, eval_err)
    avg, err = onp.azimuthal_average_array(isf._data, dist, edges, w, w2, eval_err)
    return avg, err, k, edges


@pytest.fixture(scope="module")
def isf(hip):
    img = ar1_speckle(40, 48, 40, sigma=1.5, diffusivity=0.3, mean=900, std=120, seed=21)
    return fddm.ddm(img, np.arange(1, 40), core="hip", mode="fft", precision="fp64")


@pytest.mark.parametrize("bins,rng", [(10, None), (25, (0.3, 2.5)), ([0.0, 0.4, 0.41, 1.2, 2.0, 3.5], None)])
def test_matches_oracle(isf, bins, rng):
    a = fddm.azimuthal_average(isf, bins=bins, range=rng)
    avg, err, k, edges = _ref(isf, bins, rng)
    assert np.allclose(a._data, avg, rtol=1e-12, atol=0, equal_nan=True)
    assert np.allclose(a._err, err, rtol=1e-10, atol=0, equal_nan=True)
    assert np.array_equal(np.isnan(a._data), np.isnan(avg))
    assert np.allclose(a.k, k) and np.array_equal(a.bin_edges, edges)
    assert a.shape == (len(edges), 39) and np.array_equal(a.tau, isf.tau)


def test_mask_weights_and_no_error(isf):
    rng = np.random.default_rng(3)
    mask = rng.random(isf.shape[1:]) > 0.3
    weights = rng.random(isf.shape[1:]) * 2
    a = fddm.azimuthal_average(isf, bins=12, mask=mask, weights=weights, eval_err=False)
    avg, _, _, _ = _ref(isf, 12, None, mask, weights, eval_err=False)
    assert a._err is None and a.err is None
    assert np.allclose(a._data, avg, rtol=1e-12, atol=0, equal_nan=True)


def test_device_resident_float32_result(hip):
    img = ar1_speckle(30, 64, 64, seed=5)
    lags = np.arange(1, 30)
    plan = hip.Plan("fft", img.dtype, img.shape, lags, 64, 64, 60, None, "fp32c", np.float32)
    d_img = hip.DeviceBuffer.from_array(img)
    d_out = hip.DeviceBuffer(plan.output_size())
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    plan.execute(d_img, d_out, d_ws)
    host = d_out.to_array(plan.out_shape, np.float32)
    kx = 2 * np.pi * np.fft.fftfreq(64)[:33]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(64))
    a = azimuthal_average_device(d_out, plan.out_shape, kx, ky, lags, 64, np.float32, bins=16)
    iso = fddm.ImageStructureFunction(host.astype(np.float64), kx, ky, 64, 64, lags.astype(float))
    avg, err, _, _ = _ref(iso, 16)
    assert np.allclose(a._data, avg, rtol=1e-12, atol=0, equal_nan=True)
    assert np.allclose(a._err, err, rtol=1e-9, atol=0, equal_nan=True)
    plan.destroy()
    for b in (d_img, d_out, d_ws):
        b.free()


@pytest.mark.parametrize("bins", [300, 600, 1000])
def test_many_bins(hip, bins):
    """Wide bin counts take fewer rows per workgroup (8, 4, 2, 1 by the LDS
    budget); a 128 x 128 result with 1000 bins also leaves bins empty."""
    img = ar1_speckle(21, 128, 128, sigma=1.5, diffusivity=0.3, mean=900, std=120, seed=8)
    iso = fddm.ddm(img, np.arange(1, 21), core="hip", mode="fft", precision="fp64")
    a = fddm.azimuthal_average(iso, bins=bins)
    avg, err, k, _ = _ref(iso, bins)
    assert np.allclose(a._data, avg, rtol=1e-12, atol=0, equal_nan=True)
    assert np.allclose(a._err, err, rtol=1e-10, atol=0, equal_nan=True)
    assert np.array_equal(np.isnan(a._data), np.isnan(avg))


def test_masked_nan_pixels_add_nothing(isf):
    """NaN values under the mask leave every bin finite, as in the reference,
    where masked pixels are skipped."""
    data = isf._data.copy()
    mask = np.ones(isf.shape[1:], dtype=bool)
    mask[5:9, 3:12] = False
    data[:, 5:9, 3:12] = np.nan
    bad = fddm.ImageStructureFunction(data, isf.kx, isf.ky, isf.width, isf.height, isf.tau)
    a = fddm.azimuthal_average(bad, bins=10, mask=mask)
    avg, err, _, _ = _ref(isf, 10, None, mask)
    assert np.allclose(a._data, avg, rtol=1e-12, atol=0, equal_nan=True)
    assert np.allclose(a._err, err, rtol=1e-10, atol=0, equal_nan=True)
    assert np.array_equal(np.isnan(a._data), np.isnan(avg))


def test_repeated_calls_reuse_maps(isf):
    """A second call with the same geometry returns the same values (cached
    bin maps, reused device scratch)."""
    a = fddm.azimuthal_average(isf, bins=10)
    b = fddm.azimuthal_average(isf, bins=10)
    assert np.array_equal(a._data, b._data, equal_nan=True) and np.array_equal(a._err, b._err, equal_nan=True)


def test_rejects_bad_maps(hip):
    d = hip.DeviceBuffer(8 * 3 * 4 * 2)
    w = np.ones(8)
    with pytest.raises(RuntimeError, match="decrease"):
        fddm.azimuthalaverage._reduce(d.ptr, False, 3, 2, 4, np.array([0, 1, 0, 1, 0, 0, 1, 1]), w, w, 2, True)
    with pytest.raises(RuntimeError, match="outside"):
        fddm.azimuthalaverage._reduce(d.ptr, False, 3, 2, 4, np.array([0, 1, 2, 3, 0, 0, 1, 1]), w, w, 2, True)
    d.free()
