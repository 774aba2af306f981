"""Request server (continuous batching): results are the same alone and in a
mix of concurrent requests, admission follows the HBM footprint, errors are
reported.  The kernels use no atomics and every request has its own buffers,
so "the same" is bit-for-bit."""

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _mix():
    """Eight requests of different modes, shapes, lengths and dtypes."""
    reqs = []
    shapes = [(60, 32, 32), (150, 48, 40), (35, 64, 64), (200, 30, 25), (90, 128, 128), (20, 16, 16),
              (300, 40, 48), (75, 96, 64)]
    for k, (L, H, W) in enumerate(shapes):
        dtype = np.uint16 if k % 3 else np.uint8
        img = ar1_speckle(L, H, W, sigma=1.5, diffusivity=0.2, mean=120 if dtype == np.uint8 else 900,
                          std=25 if dtype == np.uint8 else 130, dtype=dtype, seed=100 + k)
        mode = "diff" if k % 4 == 3 else "fft"
        lags = np.arange(1, L) if mode == "fft" else np.unique(np.geomspace(1, L - 1, 20).astype(int))
        prec = "fp64" if k == 5 else "fp32c"
        reqs.append(dict(mode=mode, img=img, lags=lags, nx=W, ny=H, nt=2 * L, precision=prec))
    return reqs


def _run_alone(srv, r):
    rid = srv.submit(r["mode"], r["img"], r["lags"], r["nx"], r["ny"], r["nt"], precision=r["precision"])
    srv.wait(rid, 60.0)
    out = srv.fetch(rid)
    srv.release(rid)
    return out


def test_results_are_identical_alone_and_in_the_mix(hip):
    reqs = _mix()
    srv = hip.Server(max_slots=4, hbm_bytes=2048 * MiB)
    try:
        alone = [_run_alone(srv, r) for r in reqs]
        ids = [srv.submit(r["mode"], r["img"], r["lags"], r["nx"], r["ny"], r["nt"], precision=r["precision"])
               for r in reqs]
        for rid in ids:
            srv.wait(rid, 60.0)
        mixed = [srv.fetch(rid) for rid in ids]
        times = [srv.status(rid)[1] for rid in ids]
        for rid in ids:
            srv.release(rid)
        st = srv.stats()
    finally:
        srv.close()
    for a, m in zip(alone, mixed):
        assert np.array_equal(a, m)
    # the mix really overlapped: some request was admitted before another finished
    admits = sorted(t[1] for t in times)
    dones = sorted(t[2] for t in times)
    assert admits[1] < dones[0]
    assert st["pending"] == 0 and st["running"] == 0 and st["arena_used"] == 0
    # and the server results are the direct entry point's results
    r = reqs[0]
    direct = hip.ddm_fft(r["img"], r["lags"], r["nx"], r["ny"], r["nt"], None, precision=r["precision"])
    assert np.array_equal(direct, alone[0])
    ref = onp.ddm_fft(r["img"], r["lags"], r["nx"], r["ny"], r["nt"], subtract_mean=True)
    assert np.isclose(alone[0], ref, rtol=1e-3, atol=0).all()


def test_device_resident_input_matches_host_input(hip):
    img = ar1_speckle(120, 40, 36, seed=7)
    lags = np.arange(1, 120)
    d = hip.DeviceBuffer.from_array(img)
    srv = hip.Server(max_slots=2, hbm_bytes=512 * MiB)
    try:
        a = srv.submit("fft", img, lags, 36, 40, 240)
        b = srv.submit("fft", d, lags, 36, 40, 240, img_on_device=True, shape=img.shape, dtype=img.dtype)
        srv.wait(a, 60.0)
        srv.wait(b, 60.0)
        assert np.array_equal(srv.fetch(a), srv.fetch(b))
        srv.release(a)
        srv.release(b)
    finally:
        srv.close()
        d.free()


def test_staged_host_inputs_match_device_inputs(hip):
    """Host inputs of 8 MiB and more cross through the slot's two pinned
    buffers in 32 MiB pieces (here two pieces, the second partial, then one
    partial piece of an odd-sized uint8 stack on the same slot); each result
    equals the one computed from the same frames already on the device, and
    the fetched result (staged copy) equals the same plan's output copied
    back by the runtime."""
    stacks = [ar1_speckle(700, 160, 192, seed=11),
              ar1_speckle(401, 131, 197, mean=120.0, std=25.0, dtype=np.uint8, seed=12, clip=(0, 255))]
    srv = hip.Server(max_slots=1, hbm_bytes=2048 * MiB)
    bufs = []
    try:
        for img in stacks:
            L, H, W = img.shape
            print(f"staged host input {img.shape} {img.dtype} {img.nbytes} bytes", flush=True)
            lags = np.arange(1, L)
            nx, ny, nt = next_fast_len(W, "hip"), next_fast_len(H, "hip"), next_fast_len(2 * L, "hip")
            d = hip.DeviceBuffer.from_array(img)
            bufs.append(d)
            a = srv.submit("fft", img, lags, nx, ny, nt, precision="fp32c")
            b = srv.submit("fft", d, lags, nx, ny, nt, precision="fp32c", img_on_device=True, shape=img.shape,
                           dtype=img.dtype)
            srv.wait(a, 60.0)
            srv.wait(b, 60.0)
            assert srv.status(a)[0] == srv.status(b)[0]
            got = srv.fetch(a)  # results of 8 MiB and more leave through the pinned staging pair
            assert np.array_equal(got, srv.fetch(b))
            srv.release(a)
            srv.release(b)
            # the same plan run directly, its output copied back by the runtime
            plan = hip.Plan("fft", img.dtype, img.shape, lags, nx, ny, nt, None, "fp32c")
            d_out, d_ws = hip.DeviceBuffer(plan.output_size()), hip.DeviceBuffer(plan.workspace_size())
            bufs += [d_out, d_ws]
            plan.execute(d, d_out, d_ws)
            hip.synchronize()
            ref = d_out.to_array(plan.out_shape, np.float64)
            plan.destroy()
            assert got.nbytes >= 8 * MiB
            assert np.array_equal(got, ref)
    finally:
        srv.close()
        for d in bufs:
            d.free()


def test_admission_follows_hbm_footprint(hip):
    """Two requests that only fit one at a time run one after the other; a
    small one that fits beside the first is admitted before it finishes."""
    big = np.random.default_rng(1).integers(0, 4000, (1200, 256, 256), dtype=np.uint16)
    small = ar1_speckle(30, 16, 16, seed=3)
    lb, ls = np.arange(1, 1200), np.arange(1, 30)
    probe = hip.Plan("fft", big.dtype, big.shape, lb, 256, 256, 2400, None, "fp32c")
    foot = probe.workspace_size() + probe.output_size() + big.nbytes
    probe.destroy()
    srv = hip.Server(max_slots=4, hbm_bytes=int(foot * 1.5) + 64 * MiB, aging_s=30.0)
    try:
        a = srv.submit("fft", big, lb, 256, 256, 2400)
        b = srv.submit("fft", big, lb, 256, 256, 2400)
        c = srv.submit("fft", small, ls, 16, 16, 60)
        for rid in (a, b, c):
            srv.wait(rid, 60.0)
        (_, ta), (_, tb), (_, tc) = srv.status(a), srv.status(b), srv.status(c)
        assert tb[1] >= ta[2]          # b waited for a's memory
        assert tc[1] < ta[2]           # c passed b and ran beside a
        assert np.array_equal(srv.fetch(a), srv.fetch(b))
        order = []
        while True:
            rid = srv.next_done(0.0)
            if rid == 0:
                break
            order.append(rid)
        assert sorted(order) == sorted([a, b, c]) and order.index(c) < order.index(b)
        for rid in (a, b, c):
            srv.release(rid)
    finally:
        srv.close()


def test_server_errors(hip):
    srv = hip.Server(max_slots=1, hbm_bytes=64 * MiB)
    try:
        img = np.zeros((10, 8, 8), np.uint16)
        with pytest.raises(RuntimeError, match="outside"):
            srv.submit("fft", img, [12], 8, 8, 20)
        big = np.zeros((500, 128, 128), np.uint16)
        with pytest.raises(RuntimeError, match="arena"):
            srv.submit("fft", big, np.arange(1, 500), 128, 128, 1000)
        with pytest.raises(RuntimeError):
            srv.release(12345)
        rid = srv.submit("diff", img, [1, 2], 8, 8)
        srv.wait(rid, 30.0)
        assert srv.status(rid)[0] == srv.DONE
        srv.release(rid)
    finally:
        srv.close()
    with pytest.raises(RuntimeError, match="max_slots"):
        hip.Server(max_slots=0)
