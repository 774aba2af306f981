"""fddm_plan_execute_static (the execute variant with the library's static
per-device workspace, which allocates nothing) against fddm_plan_execute with
a caller workspace, through the C ABI: bit-identical results on two calls in
a row, the second with a smaller input, in fft and diff mode; plans whose
workspace exceeds the static array are refused with an error."""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from fastddm_amd.window import blackman_harris

pytestmark = pytest.mark.gpu

CASES = [
    # (mode, precision, (L, H, W), window): the second call of each pair is smaller
    ("fft", "fp32c", (200, 96, 128), True),
    ("fft", "fp32c", (60, 64, 48), False),
    ("diff", "fp64", (120, 80, 64), False),
    ("diff", "fp64", (30, 40, 40), True),
]


def _plan(hip, mode, precision, shape, window):
    L, H, W = shape
    win = blackman_harris((H, W)) if window else None
    lags = np.arange(1, L) if mode == "fft" else np.unique(np.geomspace(1, L - 1, 20).astype(int))
    nt = next_fast_len(2 * L, "hip") if mode == "fft" else 0
    return hip.Plan(mode, np.uint16, shape, lags, W, H, nt, win, precision)


def test_static_workspace_equals_caller_workspace(hip):
    for k, (mode, precision, shape, window) in enumerate(CASES):
        print(f"static scratch call {k}: {mode} {precision} {shape}", flush=True)
        img = ar1_speckle(*shape, sigma=1.5, diffusivity=0.3, mean=1000.0, std=150.0, seed=50 + k)
        plan = _plan(hip, mode, precision, shape, window)
        assert plan.workspace_size() <= hip.Plan.STATIC_SCRATCH_BYTES
        d_img = hip.DeviceBuffer.from_array(img)
        d_ref = hip.DeviceBuffer(plan.output_size())
        d_out = hip.DeviceBuffer(plan.output_size())
        d_ws = hip.DeviceBuffer(plan.workspace_size())
        plan.execute(d_img, d_ref, d_ws)
        plan.execute_static(d_img, d_out)
        ref = d_ref.to_array(plan.out_shape, np.float64)
        got = d_out.to_array(plan.out_shape, np.float64)
        assert np.array_equal(got, ref), (k, np.max(np.abs(got - ref)))
        for b in (d_img, d_ref, d_out, d_ws):
            b.free()
        plan.destroy()


def test_static_workspace_limit(hip):
    shape = (2000, 512, 512)  # about 4.3 GB of workspace
    plan = _plan(hip, "fft", "fp32c", shape, False)
    assert plan.workspace_size() > hip.Plan.STATIC_SCRATCH_BYTES
    d_img = hip.DeviceBuffer(16)
    d_out = hip.DeviceBuffer(16)
    # straight to the C ABI: the library refuses before it touches any buffer
    with pytest.raises(RuntimeError, match="static scratch"):
        hip._call("fddm_plan_execute_static", plan.handle, d_img.ptr, d_out.ptr, None)
    d_img.free()
    d_out.free()
    plan.destroy()
