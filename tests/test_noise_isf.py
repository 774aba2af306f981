"""Noise estimators and the intermediate-scattering-function estimate:
closed forms (exact quadratics, known minima and means, a model D(q, dt)
whose ISF is an exponential) and the per-pixel loop form of the batched fit."""

import numpy as np
import pytest

from fastddm_amd import AzimuthalAverage, ImageStructureFunction
from fastddm_amd.intermediatescatteringfunction import azavg2isf_estimate
from fastddm_amd.noise_est import estimate_camera_noise


def _isf(data, width=6):
    nl, ny, nxh = data.shape[0] - 2, data.shape[1], data.shape[2]
    kx = 2 * np.pi * np.fft.fftfreq(width)[:nxh]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny))
    return ImageStructureFunction(data, kx, ky, width, ny, np.arange(1, nl + 1, dtype=float))


def _aa(data, err=None, k=None):
    nb = data.shape[0]
    k = np.linspace(0.1, 2.0, nb) if k is None else k
    return AzimuthalAverage(data, err, k, np.arange(1, data.shape[1] - 1, dtype=float), np.linspace(0, 2, nb))


def test_polyfit_recovers_the_constant_of_exact_quadratics():
    tau = np.arange(1, 9, dtype=float)
    B = np.array([3.0, 0.5, 7.0])
    d = B[:, None] + 0.3 * tau + 0.02 * tau ** 2
    aa = _aa(np.concatenate([d, np.ones((3, 2))], axis=1), err=np.full((3, 10), 0.1))
    noise, unc = estimate_camera_noise(aa, "polyfit", num_points=6)
    assert np.allclose(noise, B, rtol=1e-9)
    ny, nxh = 4, 4
    dq = (np.arange(ny * nxh).reshape(ny, nxh) * 0.25 + 1.0)
    data = dq[None] + (0.4 * tau + 0.01 * tau ** 2)[:, None, None]
    isf = _isf(np.concatenate([data, np.ones((2, ny, nxh))]))
    noise, unc = estimate_camera_noise(isf, "polyfit")
    assert np.allclose(noise, dq, rtol=1e-9) and np.allclose(unc, 0, atol=1e-9)


def test_batched_polyfit_equals_per_pixel_fits():
    rng = np.random.default_rng(1)
    data = rng.random((9, 3, 4)) + 1
    data[0, 1, 2] = np.nan
    isf = _isf(data)
    noise, unc = estimate_camera_noise(isf, "polyfit", num_points=5)
    for y in range(3):
        for x in range(4):
            if np.isnan(data[0, y, x]):
                assert np.isnan(noise[y, x]) and np.isnan(unc[y, x])
                continue
            p, cov = np.polyfit(isf.tau[:5], data[:5, y, x], deg=2, cov=True)
            assert np.isclose(noise[y, x], p[-1]) and np.isclose(unc[y, x], np.sqrt(cov[-1, -1]))


def test_min_mean_and_zero_modes():
    rng = np.random.default_rng(2)
    data = rng.random((6, 4, 4)) + 1
    data[0, 2, 3] = 0.01
    isf = _isf(data)
    assert np.all(estimate_camera_noise(isf, "min")[0] == 0.01)
    assert np.all(estimate_camera_noise(isf, "zero")[0] == 0)
    # mask: the position inside the masked selection maps through the (ky, kx) grid
    mask = np.ones((4, 4), bool)
    mask[0, 0] = False
    i = np.nanargmin(data[0, mask])
    assert np.all(estimate_camera_noise(isf, "min", mask=mask)[0] == data[0, i // 4, i % 4])
    # high_q / power_spec / var over |k| in [k_min, k_max]
    KX, KY = np.meshgrid(isf.kx, isf.ky)
    k = np.hypot(KX, KY)
    sel = (k >= 1.0) & (k <= 2.5)
    assert np.isclose(estimate_camera_noise(isf, "high_q", k_min=1.0, k_max=2.5)[0][0, 0], data[:4, sel].mean())
    assert np.isclose(estimate_camera_noise(isf, "power_spec", k_min=1.0, k_max=2.5)[0][0, 0],
                      2 * data[4][sel].mean())
    with pytest.warns(UserWarning):
        v = estimate_camera_noise(isf, "var", k_min=2.5, k_max=1.0)[0][0, 0]
    assert np.isclose(v, 2 * data[5][sel].mean())
    aa = _aa(rng.random((5, 6)), err=rng.random((5, 6)))
    n, u = estimate_camera_noise(aa, "high_q", k_min=0.5, k_max=1.5)
    s = (aa.k >= 0.5) & (aa.k <= 1.5)
    assert np.isclose(n[0], aa.data[s].mean())
    assert np.isclose(u[0], np.sqrt((aa.err[s] ** 2).sum()) / aa.err[s].size)
    with pytest.raises(RuntimeError, match="Unknown mode"):
        estimate_camera_noise(aa, "bogus")
    with pytest.raises(TypeError):
        estimate_camera_noise(np.zeros(3))


def test_isf_estimate_closed_form():
    tau = np.arange(1, 11, dtype=float)
    gamma = np.array([0.1, 0.3, 1.0])
    A, B = np.array([10.0, 5.0, 2.0]), np.array([1.0, 0.5, 0.2])
    d = B[:, None] + A[:, None] * (1 - np.exp(-gamma[:, None] * tau))
    var = (A + B) / 2
    aa = _aa(np.concatenate([d, np.zeros((3, 1)), var[:, None]], axis=1))
    isf = azavg2isf_estimate(aa, noise_est="custom", noise=B, plateau_est="var")
    assert np.allclose(isf.data, np.exp(-gamma[:, None] * tau))
    # propagated uncertainty, element by element (sigma_D = D, sigma_B = B, sigma_(A+B) = A+B)
    y, ApB = d[1, 4], A[1] + B[1]
    s2 = ((y - B[1]) / A[1] ** 2 * ApB) ** 2 + ((ApB - y) / A[1] ** 2 * B[1]) ** 2 + (y / A[1]) ** 2
    assert np.isclose(isf.err[1, 4], np.sqrt(s2))
    assert isf.shape == (3, 10) and np.array_equal(isf.tau, tau)
    with pytest.raises(RuntimeError, match="plateau_est"):
        azavg2isf_estimate(aa, noise_est="zero", plateau_est="nope")
    with pytest.raises(RuntimeError, match="noise"):
        azavg2isf_estimate(aa, noise_est="custom", noise=np.ones(5))
