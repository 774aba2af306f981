"""Pins the CPU oracles (oracle/oracle_np.py, oracle/ddm_port.c) with checks that
do not retype their formulas: brute-force DFTs and direct sums, closed forms,
invariants, the fft == diff identity, and two independent implementations
against each other.  CPU only."""

import numpy as np
import pytest

from oracle import oracle_np as onp

RNG = np.random.default_rng(1234)


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


# ---------------------------------------------------------------- FFT engines
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 11, 12, 13, 16, 25, 30, 49, 64, 77, 96,
                               121, 143, 150, 210, 343, 384, 480, 512, 1000])
def test_c_fft_matches_numpy(oracle_lib, n):
    x = RNG.standard_normal((3, n)) + 1j * RNG.standard_normal((3, n))
    assert rel(oracle_lib.fft(x), np.fft.fft(x)) < 1e-13


@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 13, 18])
def test_c_fft_matches_brute_force_dft(oracle_lib, n):
    x = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    k = np.arange(n)
    dft = np.exp(-2j * np.pi * np.outer(k, k) / n) @ x
    assert rel(oracle_lib.fft(x), dft) < 1e-13


@pytest.mark.parametrize("shape,nx,ny", [((5, 7), 7, 5), ((5, 7), 8, 6), ((6, 4), 5, 9)])
def test_numpy_spectra_match_brute_force_dft2(shape, nx, ny):
    H, W = shape
    img = RNG.integers(0, 100, (2, H, W))
    win = RNG.random((H, W))
    F = onp.spectra(img, nx, ny, win)
    for t in range(2):
        assert rel(F[t], onp.brute_force_dft2(img[t] * win, nx, ny)) < 1e-12


DTYPES = [np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64, np.uint64, np.float32, np.float64]


@pytest.mark.parametrize("dtype", DTYPES)
def test_c_rfft2_matches_numpy_all_dtypes(oracle_lib, dtype):
    img = RNG.integers(0, 200, (4, 9, 12)).astype(dtype)
    win = RNG.random((9, 12))
    for nx, ny in [(12, 9), (15, 10), (16, 16)]:
        a = oracle_lib.rfft2(img, nx, ny, win)
        b = onp.spectra(img, nx, ny, win)
        assert rel(a, b) < 1e-13


# ------------------------------------------------------- structure function
def _stack(L=24, H=10, W=12, seed=0):
    r = np.random.default_rng(seed)
    return r.integers(0, 4000, (L, H, W)).astype(np.uint16)


def test_fft_diff_and_brute_force_agree():
    img = _stack()
    lags = [1, 2, 5, 11, 23]
    nx, ny, nt = 12, 10, 48
    F = onp.spectra(img, nx, ny)
    fft = onp.ddm_fft(img, lags, nx, ny, nt)
    dif = onp.ddm_diff(img, lags, nx, ny)
    assert rel(fft, dif) < 1e-12
    for i, lag in enumerate(lags):
        bf = np.fft.fftshift(onp.brute_force_D(F, lag), axes=-2)
        assert rel(fft[i], bf) < 1e-12


def test_brute_force_ps_var():
    img = _stack(L=9)
    F = onp.spectra(img, 12, 10)
    out = onp.ddm_diff(img, [1], 12, 10)
    ps = np.zeros(F.shape[1:])
    mean = np.zeros(F.shape[1:], dtype=complex)
    for t in range(F.shape[0]):
        ps += np.abs(F[t]) ** 2
        mean += F[t]
    ps /= F.shape[0]
    mean /= F.shape[0]
    var = np.zeros(F.shape[1:])
    for t in range(F.shape[0]):
        var += np.abs(F[t] - mean) ** 2
    var /= F.shape[0]
    assert rel(out[-2], np.fft.fftshift(ps, axes=-2)) < 1e-13
    assert rel(out[-1], np.fft.fftshift(var, axes=-2)) < 1e-12


@pytest.mark.parametrize("nt", [47, 48, 64, 75, 100])
def test_fft_mode_independent_of_nt(nt):
    img = _stack()
    ref = onp.ddm_fft(img, np.arange(1, 24), 12, 10, 2 * 24)
    assert rel(onp.ddm_fft(img, np.arange(1, 24), 12, 10, nt), ref) < 1e-11


def test_mean_subtraction_leaves_d_unchanged():
    img = _stack()
    a = onp.ddm_fft(img, np.arange(1, 24), 12, 10, 48)
    b = onp.ddm_fft(img, np.arange(1, 24), 12, 10, 48, subtract_mean=True)
    assert rel(a[:-2], b[:-2]) < 1e-10


def test_static_image_added_leaves_d_unchanged():
    img = _stack().astype(np.float64)
    static = RNG.random(img.shape[1:]) * 1e3
    a = onp.ddm_diff(img, [1, 3, 7], 12, 10)
    b = onp.ddm_diff(img + static, [1, 3, 7], 12, 10)
    assert rel(a[:-2], b[:-2]) < 1e-11
    assert rel(a[-1], b[-1]) < 1e-9  # variance too


def test_constant_stack_closed_form():
    L, H, W, nx, ny, c = 6, 5, 8, 8, 6, 7.0
    img = np.full((L, H, W), c)
    out = onp.ddm_fft(img, [1, 2, 5], nx, ny, 12)
    assert np.max(np.abs(out[:-2])) < 1e-9
    assert np.max(np.abs(out[-1])) < 1e-9
    ps = out[-2]
    # all power at q = 0, which sits in row ny//2 after the shift
    f0 = c * H * W / np.sqrt(nx * ny)
    assert abs(ps[ny // 2, 0] - f0**2) < 1e-9 * f0**2
    H0, W0 = H, W
    x = np.zeros((ny, nx))
    x[:H0, :W0] = c
    full = np.abs(np.fft.rfft2(x)) ** 2 / (nx * ny)
    assert rel(ps, np.fft.fftshift(full, axes=-2)) < 1e-12


def test_impulse_gives_flat_power_spectrum():
    img = np.zeros((3, 8, 8))
    img[:, 3, 5] = 4.0
    out = onp.ddm_diff(img, [1, 2], 8, 8)
    assert np.allclose(out[-2], 16.0 / 64.0, rtol=1e-12)
    assert np.max(np.abs(out[:-2])) < 1e-12


def test_temporal_plane_wave_closed_form():
    # I = 10 + cos(2 pi k0 x / W + w t): only q = (0, +-k0) fluctuate, with
    # |F|^2 = (H W / 2)^2 / (W H), so D(dt) = 2 |F|^2 (1 - cos w dt).
    L, H, W, k0, w = 30, 8, 16, 3, 0.37
    t = np.arange(L)[:, None, None]
    x = np.arange(W)[None, None, :]
    img = 10 + np.cos(2 * np.pi * k0 * x / W + w * t) * np.ones((1, H, 1))
    lags = np.array([1, 2, 4, 9, 20])
    out = onp.ddm_fft(img, lags, W, H, 64)
    amp2 = (H * W / 2) ** 2 / (W * H)
    row0 = H // 2  # ky = 0 after the shift
    for i, dt in enumerate(lags):
        expect = 2 * amp2 * (1 - np.cos(w * dt))
        assert abs(out[i, row0, k0] - expect) < 1e-9 * amp2
        others = np.delete(out[i].ravel(), row0 * (W // 2 + 1) + k0)
        assert np.max(np.abs(others)) < 1e-9 * amp2


def test_output_shift_is_fftshift_along_y():
    img = _stack(L=5, H=7, W=6)
    F = onp.spectra(img, 6, 7)
    out = onp.ddm_diff(img, [1], 6, 7)
    ps = np.mean(np.abs(F) ** 2, axis=0)
    for i in range(7):
        assert np.allclose(out[-2][i], ps[(i - 7 // 2) % 7])


# ------------------------------------------- the C restatement vs numpy oracle
@pytest.mark.parametrize("shape,nx,ny,window", [
    ((20, 10, 12), 12, 10, False),
    ((20, 9, 11), 11, 9, True),
    ((17, 9, 11), 12, 10, True),
    ((33, 16, 16), 16, 16, False),
])
def test_c_port_matches_numpy_oracle(oracle_lib, shape, nx, ny, window):
    L, H, W = shape
    img = np.random.default_rng(5).integers(0, 3000, shape).astype(np.uint16)
    win = np.random.default_rng(6).random((H, W)) if window else None
    lags = [1, 2, 3, 7, L - 1]
    nt = 2 * L + 3
    ref_fft = onp.ddm_fft(img, lags, nx, ny, nt, win)
    assert rel(oracle_lib.ddm_fft(img, lags, nx, ny, nt, win, chunk_size=7), ref_fft) < 1e-12
    assert rel(oracle_lib.ddm_fft(img, lags, nx, ny, nt, win, chunk_size=ny * (nx // 2 + 1)), ref_fft) < 1e-12
    assert rel(oracle_lib.ddm_diff(img, lags, nx, ny, win), onp.ddm_diff(img, lags, nx, ny, win)) < 1e-12


def test_c_port_threads_do_not_change_results(oracle_lib):
    img = _stack(L=16)
    a = oracle_lib.ddm_fft(img, [1, 4, 15], 12, 10, 32, threads=1)
    b = oracle_lib.ddm_fft(img, [1, 4, 15], 12, 10, 32, threads=4)
    assert np.array_equal(a, b)


def test_reference_cross_core_relation_on_confocal(oracle_lib, confocal):
    """The reference's own equivalence test (tests/test_ddm.py:53-74): every
    core matches the cpp fft core with np.isclose defaults on the first 20
    confocal frames, lags 1..9.  Here: C restatement fft (baseline) vs its diff
    mode and vs the numpy oracle in both modes."""
    lags = np.arange(1, 10)
    nx = ny = 256
    nt = 64  # next_fast_len(2 * 20) with the 7-smooth rule is 40; any nt >= 39 works
    base = oracle_lib.ddm_fft(confocal, lags, nx, ny, 40)
    for other in (
        oracle_lib.ddm_diff(confocal, lags, nx, ny),
        onp.ddm_fft(confocal, lags, nx, ny, nt),
        onp.ddm_diff(confocal, lags, nx, ny),
    ):
        assert np.isclose(other, base).all()


def test_planted_mistake_is_caught():
    """Drop the tail term of the two-ended sum in a copy of the fft formula:
    the fft == diff identity must reject it."""
    img = _stack()
    lags = [1, 4, 9]
    F = onp.spectra(img, 12, 10)
    L = F.shape[0]
    sq = np.abs(F) ** 2
    cs = np.cumsum(sq, axis=0)
    ac = onp.autocorrelation(F, 48)
    good = onp.ddm_diff(img, lags, 12, 10)
    for i, lag in enumerate(lags):
        bad = np.fft.fftshift((cs[-(lag + 1)] - 2 * ac[lag]) / (L - lag), axes=-2)
        assert rel(bad, good[i]) > 1e-3
