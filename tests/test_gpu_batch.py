"""Batch executor: uneven stacks (mixed-radix image sizes, different lengths,
both modes) in one graph launch.  Every replay of the graph, with eager work
and synchronisation between replays, equals the eager execution of the same
plans bit for bit, and the eager results match the oracle."""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu

STACKS = [
    # (L, H, W, mode, precision)
    (60, 96, 96, "fft", "fp32c"),
    (45, 150, 200, "fft", "fp32c"),
    (130, 48, 40, "fft", "fp32c"),
    (33, 30, 42, "diff", "fp32c"),
    (90, 36, 60, "fft", "fp64"),
    (20, 50, 70, "diff", "fp64"),
    # power-of-two rows of 512: the persistent row kernel, whose FP64 form
    # with 16-frame blocks (a tuning candidate) needs more than 64 KiB of LDS
    (34, 40, 512, "fft", "fp64"),
    # long series: the time plans that run two output columns per workgroup
    # (an even and an odd number of columns)
    (2500, 12, 10, "fft", "fp32c"),
    (2300, 9, 8, "fft", "fp64"),
]


def _items(hip):
    items, hosts = [], []
    for k, (L, H, W, mode, prec) in enumerate(STACKS):
        img = ar1_speckle(L, H, W, sigma=1.5, diffusivity=0.3, mean=900, std=120, seed=40 + k)
        lags = np.arange(1, L) if mode == "fft" else np.unique(np.geomspace(1, L - 1, 12).astype(int))
        plan = hip.Plan(mode, img.dtype, img.shape, lags, W, H, next_fast_len(2 * L, "hip"), None, prec)
        out = hip.DeviceBuffer(plan.output_size())
        plan.destroy()
        items.append(dict(mode=mode, img=hip.DeviceBuffer.from_array(img), shape=img.shape, dtype=img.dtype,
                          lags=lags, nx=W, ny=H, nt=next_fast_len(2 * L, "hip"), precision=prec, out=out))
        hosts.append((img, lags))
    return items, hosts


def _outputs(items):
    return [it["out"].to_array((len(it["lags"]) + 2, it["ny"], it["nx"] // 2 + 1), np.float64) for it in items]


def test_graph_replays_equal_eager_and_oracle(hip):
    items, hosts = _items(hip)
    batch = hip.Batch(items, autotune=True)
    try:
        tbs = batch.frames_per_block()
        assert all(t in (4, 8, 16) for t in tbs)
        stream = hip.Stream()
        batch.launch_eager(stream)
        stream.synchronize()
        eager = _outputs(items)
        for rep in range(3):
            for it in items:  # clear the outputs so each replay must rewrite them
                hip.fill_uniform(it["out"], it["out"].nbytes, np.uint8, rep, 0, 256)
            batch.launch(stream)
            stream.synchronize()
            got = _outputs(items)
            for g, e in zip(got, eager):
                assert np.array_equal(g, e), rep
            batch.launch_eager(stream)  # eager work between replays
            stream.synchronize()
        stream.destroy()
    finally:
        batch.close()
    for (img, lags), e, (L, H, W, mode, prec) in zip(hosts, eager, STACKS):
        ref = onp.ddm_fft(img, lags, W, H, 2 * L, subtract_mean=True)
        if prec == "fp64":
            assert np.allclose(e, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
        else:
            assert np.isclose(e, ref, rtol=1e-3, atol=0).all()
            assert np.median(onp.relative_error(e[:-2], ref[:-2])) <= 1e-5
    for it in items:
        it["img"].free()
        it["out"].free()


def test_batch_rejects_bad_items(hip):
    img = hip.DeviceBuffer(10 * 8 * 8 * 2)
    out = hip.DeviceBuffer(1 << 20)
    bad = dict(mode="fft", img=img, shape=(10, 8, 8), dtype=np.uint16, lags=[11], nx=8, ny=8, nt=20, out=out)
    with pytest.raises(RuntimeError, match="outside"):
        hip.Batch([bad], autotune=False)
    with pytest.raises(RuntimeError):
        hip.Batch([], autotune=False)
    img.free()
    out.free()


def test_spatial_kernel_choices(hip):
    """Without tuning every row of the non-power-of-two stacks runs on the
    mixed-radix kernel (the 512-wide one on the power-of-two kernel) and every
    column (all shorter than 256) on the generic one; tuning may only switch a
    mixed-radix kernel to the generic form."""
    items, _ = _items(hip)
    default = hip.Batch(items, autotune=False)
    tuned = hip.Batch(items, autotune=True)
    try:
        k0, k1 = default.spatial_kernels(), tuned.spatial_kernels()
        assert len(k0) == len(k1) == len(STACKS)
        assert all(rows == (W != 512) and not cols for (rows, cols), (L, H, W, m, p) in zip(k0, STACKS)), k0
        for (r0, c0), (r1, c1) in zip(k0, k1):
            assert (not r1 or r0) and (not c1 or c0)
    finally:
        default.close()
        tuned.close()
    for it in items:
        it["img"].free()
        it["out"].free()
