"""fddm_ddm_fft / fddm_ddm_diff under a device-memory budget: above it the
call runs the spatial stage over chunks of frames and the time stage over
chunks of q -- with the q-major spectra in device memory (mode 1) or, when
they do not fit either, in host memory (mode 2) -- and must give the one-shot
result bit for bit.  Budgets step down from 0.8 of the plan's workspace until
the call is refused, so both modes run with several chunks; a budget too
small for one frame block and one row of q is refused with an error."""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from fastddm_amd.window import blackman_harris

pytestmark = pytest.mark.gpu

CASES = [
    # mode, precision, (L, H, W), window, lags ('all' or count of log-spaced)
    ("fft", "fp32c", (300, 64, 96), True, "all"),
    ("fft", "fp64", (200, 50, 60), False, "all"),
    ("fft", "fp32", (130, 40, 40), False, "all"),
    ("diff", "fp32c", (150, 48, 48), True, 30),
    ("diff", "fp64", (90, 35, 21), False, 20),
    # mixed-radix column kernel (ny = 300) and rows (nx = 50) in chunks
    ("fft", "fp32c", (120, 300, 50), True, "all"),
]


def _lags(L, spec):
    if spec == "all":
        return np.arange(1, L)
    return np.unique(np.geomspace(1, L - 1, spec).astype(int))


@pytest.mark.parametrize("mode,precision,shape,window,lagspec", CASES)
def test_budget_chunks_equal_one_shot(hip, mode, precision, shape, window, lagspec):
    L, H, W = shape
    img = ar1_speckle(L, H, W, sigma=1.4, diffusivity=0.2, mean=1000.0, std=150.0, seed=L + H)
    win = blackman_harris((H, W)) if window else None
    lags = _lags(L, lagspec)
    nt = next_fast_len(2 * L, "hip")

    def run():
        if mode == "fft":
            return hip.ddm_fft(img, lags, W, H, nt, win, precision=precision)
        return hip.ddm_diff(img, lags, W, H, win, precision=precision)

    plan = hip.Plan(mode, img.dtype, shape, lags, W, H, nt if mode == "fft" else 0, win, precision)
    ws = plan.workspace_size()
    plan.destroy()
    try:
        hip.set_memory_budget(0)
        ref = run()
        assert hip.last_call_chunks()["mode"] == 0
        modes = set()
        for frac in (0.8, 0.6, 0.45, 0.35, 0.25, 0.15):
            hip.set_memory_budget(int(frac * ws))
            print(f"{mode} {precision} {shape}: budget {frac} x workspace", flush=True)
            try:
                got = run()
            except RuntimeError as exc:
                assert "Not enough space" in str(exc)
                break
            st = hip.last_call_chunks()
            assert st["mode"] in (1, 2) and (st["frame_chunks"] > 1 or st["q_chunks"] > 1), st
            assert np.array_equal(got, ref), (frac, st, np.max(np.abs(got - ref)))
            modes.add(st["mode"])
        assert modes == {1, 2}, modes  # spectra on the device, then in host memory
        hip.set_memory_budget(4096)
        with pytest.raises(RuntimeError, match="Not enough space"):
            run()
    finally:
        hip.set_memory_budget(0)


def test_cache_limit(hip):
    """Buffers above the cache limit are freed when the call returns; within
    it they stay for the next call (same result either way)."""
    L, H, W = 400, 256, 256
    img = ar1_speckle(L, H, W, seed=7)
    lags = np.arange(1, L)
    nt = next_fast_len(2 * L, "hip")
    try:
        hip.release_cache()
        assert hip.cached_bytes() == 0
        hip.set_cache_limit(0)
        a = hip.ddm_fft(img, lags, W, H, nt, None, precision="fp32c")
        assert hip.cached_bytes() == 0
        hip.set_cache_limit(-1)
        b = hip.ddm_fft(img, lags, W, H, nt, None, precision="fp32c")
        # input 52 MB, spectra 2 x 106 MB, output 106 MB stay on the device
        assert hip.cached_bytes() >= 300 << 20
        assert np.array_equal(a, b)
        hip.set_cache_limit(100 << 20)  # a lower limit frees what is cached now
        assert hip.cached_bytes() == 0
        c = hip.ddm_fft(img, lags, W, H, nt, None, precision="fp32c")
        assert hip.cached_bytes() == 0 and np.array_equal(a, c)
    finally:
        hip.set_cache_limit(-1)
        hip.release_cache()
