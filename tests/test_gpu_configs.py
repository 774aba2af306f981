"""The benchmark workloads themselves, through the C ABI, against the oracle.

* Plain method: 64 x 128^2 uint16 AR(1) speckle (sigma 2, D 0.5, mean 2000,
  std 300), lags 1..63, no window -- fp64, fp32c and fp32 against the FP64
  NumPy oracle and the C restatement of the reference cpp core.
* Diff mode: 300 x 256^2 uint8 speckle (sigma 2, D 1, mean 120, std 25,
  clipped 0..255), 50 log-spaced lags in 1..299 -- fp64 and fp32c diff mode
  against the oracle's direct differences.
* Batched uneven stacks: the 48 stacks of the batch benchmark ((H, W) cycling
  96x96, 150x200, 384x384, 480x640; frames log-uniform in [50, 1500], seed
  4), fft mode, all lags, fp32c, inputs made on the device as uniform uint16
  integers.  Graph replays equal the eager run bit for bit; every fifth
  stack (all four shapes) matches the oracle on a strided set of ky rows.
* Serving: a Poisson stream of 24 mixed requests (sizes 128^2 .. 1024^2,
  one of them 1024^2 x 4000 fft with all lags; 80 % fft / 20 % diff, fp32c)
  -- every served result equals the same analysis run alone, bit for bit.

Tolerances as in test_gpu_parity.py: fp64 rtol 1e-9 (+1e-12 x slice max);
fp32c every element np.isclose(rtol=1e-3, atol=0) and median <= 1e-5.
"""

import time

import numpy as np
import pytest

from fastddm_amd import next_fast_len
from fastddm_amd.lags import logspace_int
from fastddm_amd.speckle import CONFIGS, ar1_speckle
from oracle import oracle_np as onp

pytestmark = pytest.mark.gpu


def check_fp64(out, ref):
    for s in range(ref.shape[0]):
        floor = 1e-12 * np.max(np.abs(ref[s]))
        assert np.allclose(out[s], ref[s], rtol=1e-9, atol=floor), (s, np.max(np.abs(out[s] - ref[s])))


def check_fp32c(out, ref):
    rel = onp.relative_error(out, ref)
    assert np.isclose(out, ref, rtol=1e-3, atol=0).all(), rel.max()
    assert np.median(onp.relative_error(out[:-2], ref[:-2])) <= 1e-5


def test_config_plain_128(hip, oracle_lib):
    cfg = CONFIGS["cpu_128"]
    img = ar1_speckle(**cfg)
    L, H, W = img.shape
    lags = np.arange(1, L)
    nt = next_fast_len(2 * L, "hip")
    print(f"plain 128: {img.shape} {img.dtype}, nt {nt}", flush=True)
    ref = onp.ddm_fft(img, lags, W, H, nt, subtract_mean=True)
    ref_c = oracle_lib.ddm_fft(img, lags, W, H, nt, None, chunk_size=H * (W // 2 + 1), threads=4)
    # the two oracles agree to the reference's cross-core bar (tests/test_ddm.py:53-74);
    # the C restatement, like the cpp core, does not subtract the temporal mean, so
    # near q = 0 its FP64 cancellation is larger than the NumPy oracle's
    assert np.isclose(ref_c, ref, rtol=1e-5, atol=0).all()
    check_fp64(hip.ddm_fft(img, lags, W, H, nt, None, precision="fp64"), ref)
    check_fp32c(hip.ddm_fft(img, lags, W, H, nt, None, precision="fp32c"), ref)
    out32 = hip.ddm_fft(img, lags, W, H, nt, None, precision="fp32")
    # plain fp32 (kept for precision studies, and warned about): accurate in
    # the median, not at the low-q cells where D cancels in float
    rel = np.abs(out32 - ref) / np.abs(ref)
    assert np.median(rel[:-2]) <= 1e-4
    assert np.allclose(out32[-2], ref[-2], rtol=1e-4, atol=0)


@pytest.mark.parametrize("precision", ["fp64", "fp32c"])
def test_config_diff_256(hip, precision):
    cfg = CONFIGS["diff_256"]
    img = ar1_speckle(**cfg)
    L, H, W = img.shape
    lags = logspace_int(L - 1, num=50, endpoint=True)
    assert len(np.unique(lags)) == 50 and lags.max() == L - 1
    print(f"diff 256: {img.shape} {img.dtype}, {len(lags)} lags, {precision}", flush=True)
    ref = onp.ddm_diff(img, lags, W, H)
    out = hip.ddm_diff(img, lags, W, H, None, precision=precision)
    (check_fp64 if precision == "fp64" else check_fp32c)(out, ref)
    # fft mode on the same workload agrees (Wiener-Khinchin == direct differences)
    out_fft = hip.ddm_fft(img, lags, W, H, next_fast_len(2 * L, "hip"), None, precision=precision)
    (check_fp64 if precision == "fp64" else check_fp32c)(out_fft, ref)


def _batch_items(hip):
    rng = np.random.default_rng(4)
    shapes = [(96, 96), (150, 200), (384, 384), (480, 640)]
    n = 48
    Ls = np.exp(rng.uniform(np.log(50), np.log(1500), n)).round().astype(int)
    items = []
    for k in range(n):
        H, W = shapes[k % 4]
        L = int(Ls[k])
        img = hip.DeviceBuffer(L * H * W * 2)
        hip.fill_uniform(img, L * H * W, np.uint16, 7000 + k, 800, 1200)
        lags = np.arange(1, L)
        nt = next_fast_len(2 * L, "hip")
        out = hip.DeviceBuffer((L + 1) * H * (W // 2 + 1) * 8)
        items.append(dict(mode="fft", img=img, shape=(L, H, W), dtype=np.uint16, lags=lags, nx=W, ny=H,
                          nt=nt, precision="fp32c", out=out))
    return items


def test_config_batch_48_stacks(hip):
    items = _batch_items(hip)
    print(f"batch: {len(items)} stacks, {sum(it['shape'][0] for it in items)} frames", flush=True)
    batch = hip.Batch(items, autotune=True)
    batch.launch_eager()
    hip.synchronize()
    eager = [it["out"].to_array(((it["shape"][0] + 1) * it["ny"] * (it["nx"] // 2 + 1),), np.float64)
             for it in items]
    for it in items:  # poison the outputs: the replays must rewrite them
        hip.fill_uniform(it["out"], it["out"].nbytes, np.uint8, 1, 0, 255)
    for rep in range(3):
        batch.launch()
        hip.synchronize()
        for k, it in enumerate(items):
            got = it["out"].to_array(eager[k].shape, np.float64)
            assert np.array_equal(got, eager[k]), (rep, k)
    print("batch: graph replays == eager", flush=True)
    for k in range(0, len(items), 5):
        it = items[k]
        L, H, W = it["shape"]
        img = it["img"].to_array((L, H, W), np.uint16)
        F = onp.spectra(img, W, H)
        rows = np.arange(0, H, 7)  # unshifted ky rows
        ref = onp.structure_from_spectra(F[:, rows, :], it["lags"], 2 * L, subtract_mean=True)
        out = eager[k].reshape(L + 1, H, W // 2 + 1)[:, (rows + H // 2) % H, :]
        check_fp32c(out, ref)
        print(f"batch stack {k} ({L} x {H} x {W}) matches the oracle", flush=True)
    batch.close()
    for it in items:
        it["img"].free()
        it["out"].free()


def test_config_serving_stream_alone_equals_mix(hip):
    rng = np.random.default_rng(15)
    n = 24
    sizes = np.array([128, 256, 512, 1024])
    size = sizes[rng.integers(0, 3, n)]  # the large one is forced below
    L = np.exp(rng.uniform(np.log(100), np.log(1500), n)).round().astype(int)
    mode = np.where(rng.random(n) < 0.8, "fft", "diff")
    big = n // 3
    size[big], L[big], mode[big] = 1024, 4000, "fft"
    arrivals = np.cumsum(rng.exponential(1.0 / 200.0, n))
    stacks = {}
    for k, sz in enumerate(sizes):
        frames = 4000 if sz == 1024 else 1500
        b = hip.DeviceBuffer(frames * int(sz) * int(sz) * 2)
        hip.fill_uniform(b, frames * int(sz) * int(sz), np.uint16, 300 + k, 800, 1200)
        stacks[int(sz)] = b

    def args(i):
        sz, l = int(size[i]), int(L[i])
        lags = np.arange(1, l) if mode[i] == "fft" else logspace_int(l - 1, num=50, endpoint=True)
        nt = next_fast_len(2 * l, "hip") if mode[i] == "fft" else 0
        return sz, l, lags, nt

    print(f"serving: {n} requests, request {big} is 1024^2 x 4000 fft", flush=True)
    srv = hip.Server(max_slots=4)
    rid = {}
    t0 = time.perf_counter()
    for i in range(n):
        wait = arrivals[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        sz, l, lags, nt = args(i)
        rid[i] = srv.submit(mode[i], stacks[sz], lags, sz, sz, nt, precision="fp32c", img_on_device=True,
                            shape=(l, sz, sz), dtype=np.uint16)
    mixed = {}
    for i in range(n):
        srv.wait(rid[i], 300.0)
        mixed[i] = srv.fetch(rid[i])
        srv.release(rid[i])
    srv.close()
    print("serving: all requests served", flush=True)
    for i in range(n):
        sz, l, lags, nt = args(i)
        plan = hip.Plan(mode[i], np.uint16, (l, sz, sz), lags, sz, sz, nt, None, "fp32c")
        d_out = hip.DeviceBuffer(plan.output_size())
        d_ws = hip.DeviceBuffer(plan.workspace_size())
        plan.execute(stacks[sz], d_out, d_ws)
        alone = d_out.to_array(plan.out_shape, np.float64)
        assert np.array_equal(alone, mixed[i]), (i, mode[i], sz, l)
        plan.destroy()
        d_out.free()
        d_ws.free()
        del mixed[i]
    for b in stacks.values():
        b.free()
