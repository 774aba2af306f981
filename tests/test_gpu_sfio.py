"""A device-resident result streamed to a ``.sf.ddm`` file in small chunks
is byte-identical to the file written from the same result on the host."""

import numpy as np
import pytest

from fastddm_amd import ImageStructureFunction, sfio
from fastddm_amd.speckle import ar1_speckle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("out_dtype", [np.float64, np.float32])
def test_streamed_file_equals_host_file(hip, tmp_path, out_dtype):
    img = ar1_speckle(25, 40, 36, seed=12)
    lags = np.arange(1, 25)
    plan = hip.Plan("fft", img.dtype, img.shape, lags, 36, 40, 50, None, "fp32c", out_dtype)
    d_img = hip.DeviceBuffer.from_array(img)
    d_out = hip.DeviceBuffer(plan.output_size())
    d_ws = hip.DeviceBuffer(plan.workspace_size())
    plan.execute(d_img, d_out, d_ws)
    host = d_out.to_array(plan.out_shape, out_dtype)
    kx = (2 * np.pi * np.fft.fftfreq(36)[:19]).astype(out_dtype)
    ky = (2 * np.pi * np.fft.fftshift(np.fft.fftfreq(40))).astype(out_dtype)
    tau = lags.astype(out_dtype)
    a = ImageStructureFunction(host, kx, ky, 36, 40, tau).save(str(tmp_path / "host"))
    b = sfio.save_device_result(str(tmp_path / "dev"), d_out, plan.out_shape, kx, ky, tau, 36, 40, out_dtype,
                                chunk_bytes=8192)
    assert open(a, "rb").read() == open(b, "rb").read()
    back = sfio.load(b)
    assert np.array_equal(back._data, host)
    plan.destroy()
    for buf in (d_img, d_out, d_ws):
        buf.free()
