"""GPU parity: the HIP core (through the C ABI) against the CPU oracles on the
same seeded inputs.  Tolerances (stated per test):

* fp64: every element within rtol 1e-9 of the FP64 oracle, plus an absolute
  floor of 1e-12 x max|ref| per slice (cancellation at the smallest values);
* fp32c: the reference's single-precision bar, np.isclose(rtol=1e-3, atol=0)
  on every D / PS / VAR element (tests/test_ddm.py:116-157), and median
  relative error <= 1e-5;
* fp32 (plain): median relative error <= 1e-4 on D, PS within rtol 1e-4.
"""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu


def relerr(x, ref):
    return onp.relative_error(x, ref)


def check_fp64(out, ref):
    for s in range(ref.shape[0]):
        floor = 1e-12 * np.max(np.abs(ref[s]))
        assert np.allclose(out[s], ref[s], rtol=1e-9, atol=floor), (s, np.max(np.abs(out[s] - ref[s])))


def check_fp32c(out, ref):
    assert np.isclose(out, ref, rtol=1e-3, atol=0).all(), np.max(relerr(out, ref))
    assert np.median(relerr(out[:-2], ref[:-2])) <= 1e-5


def check_fp32(out, ref):
    assert np.median(relerr(out[:-2], ref[:-2])) <= 1e-4
    assert np.allclose(out[-2], ref[-2], rtol=1e-4, atol=0)


CHECK = {"fp64": check_fp64, "fp32c": check_fp32c, "fp32": check_fp32}


# ------------------------------------------------------------ device utilities
def test_device_utilities(hip):
    assert hip.get_num_devices() >= 1
    hip.set_device(0)
    assert hip.get_device() == 0
    assert hip.get_free_device_memory() > 0
    with pytest.raises(RuntimeError):
        hip.set_device(hip.get_num_devices() + 3)


# ------------------------------------------------------------- FFT engine
FFT_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 16, 25, 27, 32, 49, 64, 96, 125, 128, 150, 200, 243,
             256, 343, 384, 480, 512, 600, 640, 1000, 1024, 2048, 2401, 4000, 4096, 6000, 8000]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("fp32", 2e-6)])
def test_fft_engine_matches_numpy(hip, precision, tol):
    rng = np.random.default_rng(7)
    for n in FFT_SIZES:
        if precision == "fp64" and n > 8192:
            continue
        x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
        got = hip.device_fft(x, precision)
        ref = np.fft.fft(x)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        assert err < tol * max(1.0, np.log2(max(n, 2)) / 4), (n, err)


# ------------------------------------------------------ spatial spectra stage
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
@pytest.mark.parametrize("shape,nx,ny,window", [
    ((21, 30, 40), 40, 30, False),
    ((18, 35, 21), 21, 35, True),
    ((9, 30, 40), 42, 32, True),
    # power-of-two kernels (256 .. 4096): exact sizes and zero padding
    ((6, 256, 256), 256, 256, True),
    ((5, 255, 509), 512, 256, True),
    ((4, 512, 1024), 1024, 512, False),
    ((3, 1024, 256), 256, 1024, True),
    # 4-frame blocks (fp64 from ny = 1024, fp32 from ny = 2048): two pair rows per workgroup
    ((3, 1024, 1024), 1024, 1024, True),
    ((3, 2048, 256), 256, 2048, False),
])
def test_spatial_spectra_match_oracle(hip, precision, shape, nx, ny, window):
    L, H, W = shape
    img = np.random.default_rng(11).integers(0, 4000, shape).astype(np.uint16)
    win = np.random.default_rng(12).random((H, W)) if window else None
    plan = hip.Plan("fft", img.dtype, shape, [1], nx, ny, next_fast_len(2 * L, "hip"), win, precision)
    d_img = hip.DeviceBuffer.from_array(img)
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    got = plan.spectra(d_img, d_ws)  # (nq, L), q = ky * nxh + kx
    base = img.astype(np.float64)
    if precision != "fp32":
        base = base - base[0]  # the compensated paths transform I(t) - I(0)
    ref = onp.spectra(base, nx, ny, win).reshape(L, -1).T
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < (1e-13 if precision == "fp64" else 2e-6), err


@pytest.mark.parametrize("precision", ["fp32", "fp32c", "fp64"])
@pytest.mark.parametrize("shape,dtype,window", [
    ((131, 256, 256), np.uint16, True),   # 9 frame blocks x 256 rows; odd L: last pair half empty
    ((70, 512, 512), np.int16, False),    # 5 frame blocks x 512 rows, signed pixels
    ((67, 400, 512), np.uint8, True),     # 8 pixels per vector
    ((45, 300, 256), np.int32, False),    # 2 pixels per vector
    ((38, 256, 512), np.float32, True),
    ((21, 150, 256), np.uint32, True),
])
def test_persistent_rows_match_oracle(hip, precision, shape, dtype, window):
    """More (row, frame block) items than resident workgroups, so each
    workgroup of the persistent row kernel loops over items; every pixel
    size it takes."""
    L, H, W = shape
    rng = np.random.default_rng(13)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        img = rng.integers(max(info.min, -3000), min(info.max, 4000), shape, endpoint=True).astype(dtype)
    else:
        img = (rng.random(shape) * 4000).astype(dtype)
    win = np.random.default_rng(14).random((H, W)) if window else None
    plan = hip.Plan("fft", img.dtype, shape, [1], W, H, next_fast_len(2 * L, "hip"), win, precision)
    d_img = hip.DeviceBuffer.from_array(img)
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    got = plan.spectra(d_img, d_ws)
    base = img.astype(np.float64)
    if precision != "fp32":
        base = base - base[0]
    ref = onp.spectra(base, W, H, win).reshape(L, -1).T
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < (1e-13 if precision == "fp64" else 2e-6), err
    plan.destroy()
    d_img.free()
    d_ws.free()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp32c"])
@pytest.mark.parametrize("shape,dtype,nx,ny,window", [
    ((70, 150, 200), np.uint16, 200, 150, True),    # 8-byte pixel vectors
    ((45, 96, 150), np.uint8, 150, 96, False),      # row not a whole number of vectors
    ((33, 384, 384), np.int16, 384, 384, True),
    ((17, 480, 640), np.float32, 640, 480, True),   # 2-pixel vectors
    ((12, 100, 120), np.uint16, 125, 105, True),    # zero padding in x and y
    ((9, 50, 70), np.int64, 72, 56, False),         # 8-byte pixels: one per load
    ((8, 700, 36), np.uint16, 36, 700, False),      # 4 frames per block (fp64), 8 (fp32)
    ((21, 30, 768), np.uint16, 768, 30, True),      # longest row of the mixed-radix kernels
    ((5, 768, 30), np.int16, 30, 768, False),       # longest column of the short form
    ((100, 480, 640), np.uint16, 640, 480, True),   # more items than resident workgroups
    ((7, 392, 686), np.uint16, 686, 392, True),     # radix 7 in rows (2 7 7 7) and columns (8 7 7)
    ((5, 270, 343), np.int16, 343, 270, False),     # odd row length 7^3; columns 2 3 3 3 5
    # long form (769 .. 2048 points)
    ((6, 40, 1280), np.uint16, 1280, 40, True),
    ((5, 1080, 30), np.uint16, 30, 1080, False),
    ((4, 64, 1920), np.float32, 1920, 64, True),
    ((3, 20, 2000), np.int16, 2000, 20, False),
    ((3, 1000, 1000), np.uint16, 1000, 1000, True),
])
def test_mixed_radix_spectra_match_oracle(hip, precision, shape, dtype, nx, ny, window):
    """Non-power-of-two rows and columns up to 2048: the one-wave-per-transform
    mixed-radix kernels, every pixel load form and frame-block size."""
    L, H, W = shape
    rng = np.random.default_rng(L + H + W)
    if np.issubdtype(dtype, np.integer):
        info = np.iinfo(dtype)
        img = rng.integers(max(info.min, -3000), min(info.max, 4000), shape, endpoint=True).astype(dtype)
    else:
        img = (rng.random(shape) * 4000).astype(dtype)
    win = rng.random((H, W)) if window else None
    plan = hip.Plan("fft", img.dtype, shape, [1], nx, ny, next_fast_len(2 * L, "hip"), win, precision)
    d_img = hip.DeviceBuffer.from_array(img)
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    got = plan.spectra(d_img, d_ws)
    base = img.astype(np.float64)
    if precision != "fp32":
        base = base - base[0]
    ref = onp.spectra(base, nx, ny, win).reshape(L, -1).T
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < (1e-13 if precision == "fp64" else 2e-6), err
    plan.destroy()
    d_img.free()
    d_ws.free()


# ------------------------------------------------ whole path against the oracle
CASES = [
    # (L, H, W, nx, ny, window, lags)
    (40, 32, 32, 32, 32, False, "all"),
    (37, 30, 25, 25, 30, True, "all"),
    (25, 21, 27, 27, 21, False, [1, 2, 3, 10, 24]),
    (16, 24, 20, 21, 25, True, [1, 5, 15]),  # zero padding in x and y
    (64, 20, 48, 48, 20, False, "all"),
    (24, 256, 256, 256, 256, True, "all"),   # power-of-two spatial kernels
    (19, 250, 255, 256, 256, False, [1, 2, 9, 18]),
]


@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32"])
@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("mode", ["fft", "diff"])
def test_whole_path_matches_oracle(hip, mode, case, precision):
    L, H, W, nx, ny, window, lags = case
    img = ar1_speckle(L, H, W, sigma=1.5, diffusivity=0.3, mean=1000, std=150, seed=L + H)
    win = np.random.default_rng(3).random((H, W)) if window else None
    lags = np.arange(1, L) if lags == "all" else np.array(lags)
    ref = onp.ddm_fft(img, lags, nx, ny, 2 * L, win, subtract_mean=True)
    if mode == "fft":
        nt = next_fast_len(2 * L, "hip")
        out = hip.ddm_fft(img, lags, nx, ny, nt, win, precision=precision)
    else:
        out = hip.ddm_diff(img, lags, nx, ny, win, precision=precision)
    assert out.shape == ref.shape
    CHECK[precision](out, ref)


# Every power-of-two time plan (N = 256 .. 16384: three- and four-pass plans,
# 64 .. 1024 threads per transform) on a small image.
TIME_LENGTHS = [(100, 256), (200, 512), (300, 1024), (700, 2048), (1500, 4096), (3000, 8192), (5000, 16384),
                (10000, 32768)]


@pytest.mark.parametrize("precision", ["fp64", "fp32c", "fp32"])
@pytest.mark.parametrize("L,N", TIME_LENGTHS)
def test_every_time_plan_matches_oracle(hip, L, N, precision):
    assert 2 * L - 1 <= N < 4 * L - 2 or N == 256
    img = ar1_speckle(L, 12, 10, sigma=1.2, diffusivity=0.05, mean=800, std=120, seed=L)
    lags = np.unique(np.concatenate([np.arange(1, 40), np.geomspace(40, L - 1, 60).astype(int),
                                     np.arange(L - 40, L)]))
    nt = next_fast_len(2 * L, "hip")
    # fp64 from 5000 frames and fp32 from 10000 run the four-step time
    # transform through device scratch (N = 4096 x 4 or x 8)
    ref = onp.ddm_fft(img, lags, 10, 12, 2 * L, None, subtract_mean=True)
    out = hip.ddm_fft(img, lags, 10, 12, nt, None, precision=precision)
    CHECK[precision](out, ref)


@pytest.mark.parametrize("precision,L", [("fp32c", 700), ("fp32c", 1500), ("fp32", 1500), ("fp32c", 2500),
                                         ("fp32c", 4500), ("fp64", 1500), ("fp64", 2500), ("fp32", 2500)])
def test_time_plan_odd_number_of_q(hip, precision, L):
    """45 q-series (9 x 8 images): the plans that put several q in one
    workgroup (two at N = 4096, fp32c and fp32) end with a partly empty workgroup
    whose idle groups still serve the lag loop of the others; the plans that
    run two output columns one after the other (N = 8192 and 16384 fp32c,
    N = 4096 and 8192 fp64, N = 8192 fp32) end with a lone column, and with 45 values per lag row
    the column pairs are stored as two values rather than one aligned pair."""
    img = ar1_speckle(L, 9, 8, sigma=1.1, diffusivity=0.08, mean=700, std=110, seed=L + 7)
    lags = np.arange(1, L)
    ref = onp.ddm_fft(img, lags, 8, 9, 2 * L, None, subtract_mean=True)
    out = hip.ddm_fft(img, lags, 8, 9, next_fast_len(2 * L, "hip"), None, precision=precision)
    CHECK[precision](out, ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
@pytest.mark.parametrize("L", [120, 200, 1500, 5000])
def test_diff_mode_long_series(hip, L, precision):
    """Every shape of the diff kernel: four q per wave (fp32c at 120), two
    (fp64 at 120, fp32c at 200), one q per wave from LDS (fp64 at 200, fp32c
    at 1500) and from L2 (fp64 at 1500, both at 5000)."""
    img = ar1_speckle(L, 8, 9, sigma=1.0, diffusivity=0.1, mean=600, std=90, seed=L + 1)
    lags = np.unique(np.geomspace(1, L - 1, 50).astype(int))
    ref = onp.ddm_fft(img, lags, 9, 8, 2 * L, None, subtract_mean=True)
    CHECK[precision](hip.ddm_diff(img, lags, 9, 8, None, precision=precision), ref)


@pytest.mark.parametrize("dtype", [np.uint8, np.int16, np.uint16, np.int32, np.uint32, np.int64,
                                   np.uint64, np.float32, np.float64])
def test_every_input_dtype(hip, dtype):
    img = np.random.default_rng(4).integers(0, 250, (12, 16, 18)).astype(dtype)
    lags = np.arange(1, 12)
    ref = onp.ddm_fft(img, lags, 18, 16, 24, subtract_mean=True)
    check_fp64(hip.ddm_fft(img, lags, 18, 16, 24, None, precision="fp64"), ref)
    check_fp64(hip.ddm_diff(img, lags, 18, 16, None, precision="fp64"), ref)
    check_fp32c(hip.ddm_fft(img, lags, 18, 16, 24, None, precision="fp32c"), ref)


@pytest.mark.parametrize("L", [20, 2500])
def test_float32_output_dtype(hip, L):
    """Also at N = 8192, whose column pairs are stored as one 8-byte pair."""
    img = ar1_speckle(L, 16, 16, seed=9)
    lags = np.arange(1, L)
    ref = onp.ddm_fft(img, lags, 16, 16, 2 * L, subtract_mean=True)
    out = hip.ddm_fft(img, lags, 16, 16, next_fast_len(2 * L, "hip"), None, precision="fp32c",
                      out_dtype=np.float32)
    assert out.dtype == np.float32
    assert np.isclose(out, ref, rtol=1e-3, atol=0).all()


def test_no_lags_gives_power_spectrum_and_variance(hip):
    img = ar1_speckle(10, 16, 16, seed=2)
    ref = onp.ddm_diff(img, [], 16, 16)
    for mode in ("fft", "diff"):
        out = (hip.ddm_fft(img, [], 16, 16, 20, None) if mode == "fft"
               else hip.ddm_diff(img, [], 16, 16, None))
        assert out.shape == (2, 16, 9)
        check_fp64(out, ref)


@pytest.mark.parametrize("precision,L", [("fp64", 1100), ("fp32c", 2100)])
def test_no_lags_on_two_column_plans(hip, precision, L):
    """No lags on the plans that run two output columns per workgroup
    (FP64 N = 4096, fp32c N = 8192): only PS and VAR are written."""
    img = ar1_speckle(L, 9, 8, seed=L)
    ref = onp.ddm_diff(img, [], 8, 9)
    out = hip.ddm_fft(img, [], 8, 9, next_fast_len(2 * L, "hip"), None, precision=precision)
    assert out.shape == (2, 9, 5)
    if precision == "fp64":
        check_fp64(out, ref)
    else:  # PS and VAR rows only: the fp32c bar element-wise, and a tight median
        assert np.isclose(out, ref, rtol=1e-3, atol=0).all()
        assert np.median(relerr(out, ref)) <= 1e-5


def test_two_frames(hip):
    img = ar1_speckle(2, 12, 10, seed=5)
    ref = onp.ddm_diff(img, [1], 10, 12)
    check_fp64(hip.ddm_fft(img, [1], 10, 12, 4, None), ref)
    check_fp64(hip.ddm_diff(img, [1], 10, 12, None), ref)


def test_reference_cross_core_on_confocal(hip, confocal):
    """tests/test_ddm.py:53-94 of the reference: every core equals the cpp fft
    core with np.isclose defaults on confocal[:20], lags 1..9 -- here against
    the FP64 oracle, for the HIP fft and diff modes in fp64."""
    lags = np.arange(1, 10)
    ref = onp.ddm_fft(confocal, lags, 256, 256, 40)
    nt = next_fast_len(40, "hip")
    assert np.isclose(hip.ddm_fft(confocal, lags, 256, 256, nt, None), ref).all()
    assert np.isclose(hip.ddm_diff(confocal, lags, 256, 256, None), ref).all()
    # single-precision bar of the reference (rtol 1e-3, atol 0)
    assert np.isclose(hip.ddm_fft(confocal, lags, 256, 256, nt, None, precision="fp32c"), ref,
                      rtol=1e-3, atol=0).all()


# ------------------------------------------------------------ ddm() front end
def test_ddm_front_end_with_hip_core(hip):
    import fastddm_amd as fddm

    img = ar1_speckle(30, 24, 20, seed=8)
    win = fddm.window.blackman_harris(img.shape)
    lags = [0, 3, 1, 3, 29, 7]
    with pytest.warns(UserWarning):
        res = fddm.ddm(img, lags, core="hip", mode="fft", window=win)
    ref = onp.ddm_fft(img, [1, 3, 7, 29], 20, 24, 60, win, subtract_mean=True)
    check_fp64(res._data, ref)
    assert np.array_equal(res.tau, [1, 3, 7, 29])
    assert res.shape == (4, 24, 11)
    res_d = fddm.ddm(img, [1, 3, 7, 29], core="hip", mode="diff", window=win, precision="fp32c")
    check_fp32c(res_d._data, ref)
    with pytest.raises(RuntimeError):
        fddm.ddm(img, [1], core="hip", precision="fp16")
    with pytest.raises(RuntimeError):
        fddm.ddm(img, [1], core="hip", gpu_id=99)


@pytest.mark.parametrize("precision,L,size", [("fp64", 5000, 64), ("fp32c", 9000, 48), ("fp32", 9000, 48),
                                               ("fp64", 20000, 6), ("fp32c", 33000, 4)])
def test_four_step_time_transform(hip, precision, L, size):
    """Series beyond the LDS-resident plans (fp64 over 4096 frames, fp32 over
    8192): the four-step transform of length 4096 x N1 (N1 = 4 .. 16) with one
    workgroup per q; at 64 x 64 and 48 x 48 there are more q than scratch slots,
    so workgroups loop over several q.  All lags for the short images, a
    log-spaced set plus the last 40 for the large ones."""
    img = ar1_speckle(L, size, size, sigma=1.2, diffusivity=0.05, mean=800, std=120, seed=L + size)
    if size <= 8:
        lags = np.arange(1, L)
    else:
        lags = np.unique(np.concatenate([np.arange(1, 40), np.geomspace(40, L - 1, 80).astype(int),
                                         np.arange(L - 40, L)]))
    nt = next_fast_len(2 * L, "hip")
    ref = onp.ddm_fft(img, lags, size, size, 2 * L, None, subtract_mean=True)
    out = hip.ddm_fft(img, lags, size, size, nt, None, precision=precision)
    CHECK[precision](out, ref)


def test_too_long_series_is_rejected(hip):
    # beyond 2^16 points in fp64 (32768 frames)
    img = np.zeros((40000, 4, 4), np.uint8)
    with pytest.raises(RuntimeError, match="too long"):
        hip.ddm_fft(img, [1], 4, 4, 80000, None, precision="fp64")


def test_execute_refuses_misaligned_buffers(hip):
    """Input, workspace and output pointers below the kernels' vector widths
    are refused before any launch (input 8 bytes, workspace 16, output its
    element); aligned ones run."""
    import ctypes

    L, H, W = 8, 16, 16
    plan = hip.Plan("fft", np.uint16, (L, H, W), [1, 2], W, H, 16, None, "fp32c")
    d_img = hip.DeviceBuffer(L * H * W * 2 + 64)
    d_out = hip.DeviceBuffer(plan.output_size() + 64)
    d_ws = hip.DeviceBuffer(plan.workspace_size() + 64)
    err = ctypes.create_string_buffer(256)
    try:
        for img_off, ws_off, out_off in [(4, 0, 0), (2, 0, 0), (0, 8, 0), (0, 0, 4)]:
            rc = hip.lib().fddm_plan_execute(plan.handle, ctypes.c_void_p(d_img.ptr.value + img_off),
                                             ctypes.c_void_p(d_out.ptr.value + out_off),
                                             ctypes.c_void_p(d_ws.ptr.value + ws_off), None, err, 256)
            assert rc != 0 and b"aligned" in err.value, (img_off, ws_off, out_off, rc, err.value)
        rc = hip.lib().fddm_plan_execute(plan.handle, d_img.ptr, d_out.ptr, d_ws.ptr, None, err, 256)
        assert rc == 0, err.value
        hip.synchronize()
    finally:
        plan.destroy()
        for b in (d_img, d_out, d_ws):
            b.free()
