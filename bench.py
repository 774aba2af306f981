"""Benchmark of the fft-mode image structure function on one MI355X.

Workload (default): 2000 frames of 512x512 uint16 AR(1) speckle (sigma 1.5 px,
D 0.2 px^2/frame, mean 1000, std 150, Poisson noise, seed 2), lags 1..1999,
Blackman-Harris window, time FFT length 4000, float64 output.  One step is one
whole analysis with the raw stack already in device memory:
frame-0 spectra -> row pass -> column pass -> time kernel, output in device
memory.  Reported:

* value: frames/s of the fp32c path (L / seconds per step), the FP64 GPU path
  timed beside it, the ratio, and the fp32c error against the FP64 GPU result;
* per-stage times from HIP events on the execute stream, and the roofline of
  the dominant kernel from its algorithmic bytes per launch;
* cpu_baseline: the FP64 C restatement of the reference C++ core
  (oracle/ddm_port.c, single thread) on a bounded sample of the same workload,
  scaled to the full job.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--skip-cpu]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VECTOR_PEAK_TFLOPS = 78.6  # MI355X FP64 vector peak


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(L, H, W, b_in, nq, nl, c, o, window):
    """Compulsory bytes per stage: inputs read once, outputs written once.

    rows: pixels in, row spectra out (L H (W/2+1) complex); columns: row
    spectra in, q-major spectra out; time: q-major spectra in, D/PS/VAR out."""
    nxh = W // 2 + 1
    spectra = L * nq * c
    out = (nl + 2) * nq * o
    rows = L * H * W * b_in + L * H * nxh * c + (H * W * (c // 2) if window else 0)
    cols = L * H * nxh * c + spectra
    time_stage = spectra + out
    whole = L * H * W * b_in + 2 * spectra + out
    return {"rows": rows, "cols": cols, "spatial": L * H * W * b_in + spectra, "time": time_stage, "whole": whole}


def stage_fractions(st, b):
    """Algorithmic bytes / stage time / peak for the row, column and time stages."""
    fr = {}
    for key, ms_key in (("rows", "row_pass"), ("cols", "column_pass"), ("time", "time_or_diff")):
        ms = st.get(ms_key)
        if ms:
            gbs = b[key] / (ms / 1e3) / 1e9
            fr[key] = {"GB": b[key] / 1e9, "ms": ms, "GBs": gbs, "frac": gbs / HBM_PEAK_GBS}
    return fr


def _time_transform_length(L):
    n = 512
    while n < 2 * L - 1:
        n *= 2
    return n


def run_gpu(core, img, lags, nx, ny, nt, win, precision, steps, warmup, keep_output, mode="fft"):
    L, H, W = img.shape
    plan = core.Plan(mode, img.dtype, img.shape, lags, nx, ny, nt, win, precision, np.float64)
    d_img = core.DeviceBuffer.from_array(img)
    d_out = core.DeviceBuffer(plan.output_size())
    d_ws = core.DeviceBuffer(plan.workspace_size())
    stream = core.Stream()
    for _ in range(warmup):
        plan.execute(d_img, d_out, d_ws, stream)
    stream.synchronize()
    plan.timing_begin(steps)
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.execute(d_img, d_out, d_ws, stream)
    stream.synchronize()
    wall = time.perf_counter() - t0
    stage_ms, n = plan.timing_end()
    stage_avg = {k: v / max(n, 1) for k, v in stage_ms.items()}
    out = d_out.to_array(plan.out_shape, np.float64) if keep_output else None
    azavg_ms = None
    if mode == "fft" and precision == "fp32c":
        # azimuthal average (50 bins, with errors) of the device-resident result
        from fastddm_amd.azimuthalaverage import azimuthal_average_device

        kx = 2 * np.pi * np.fft.fftfreq(nx)[: nx // 2 + 1]
        ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny))
        tau = np.asarray(lags, dtype=np.float64)
        azimuthal_average_device(d_out, plan.out_shape, kx, ky, tau, nx, np.float64, bins=50)
        t = time.perf_counter()
        for _ in range(3):
            azimuthal_average_device(d_out, plan.out_shape, kx, ky, tau, nx, np.float64, bins=50)
        azavg_ms = (time.perf_counter() - t) / 3 * 1e3
    stage_avg["azimuthal_average_50_bins"] = azavg_ms
    for b in (d_img, d_out, d_ws):
        b.free()
    stream.destroy()
    plan.destroy()
    return wall / steps * 1e3, stage_avg, out


def _cpu_model():
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(img, lags, nx, ny, nt, win, full_L, nq, threads):
    """FP64 C restatement of the reference cpp core on a sample of the
    benchmark workload, scaled to the full job (frames/s)."""
    from oracle import port

    L, H, W = img.shape
    # spatial stage: 100 of the frames at full size
    fs = min(100, L)
    t0 = time.perf_counter()
    port.rfft2(img[:fs], nx, ny, win, threads=threads)
    t_spatial = (time.perf_counter() - t0) / fs * full_L
    # time stage: all frames, a 16-row crop (16 * (nx/2+1) q-series)
    crop = np.ascontiguousarray(img[:, :16, :])
    wcrop = None if win is None else np.ascontiguousarray(win[:16, :])
    nq_crop = 16 * (nx // 2 + 1)
    t0 = time.perf_counter()
    port.rfft2(crop, nx, 16, wcrop, threads=threads)
    t_crop_spatial = time.perf_counter() - t0
    t0 = time.perf_counter()
    port.ddm_fft(crop, lags, nx, 16, nt, wcrop, chunk_size=257, threads=threads)
    t_crop_total = time.perf_counter() - t0
    t_time = max(t_crop_total - t_crop_spatial, 0.0) / nq_crop * nq
    total = t_spatial + t_time
    return {
        "value": full_L / total,
        "unit": "frames/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"spatial stage timed on {fs} of {full_L} frames ({H}x{W}), time stage on all "
                   f"{full_L} frames for {nq_crop} of {nq} q-series (16-row crop); both scaled "
                   f"linearly to the full job; est. {total:.1f} s per analysis"),
    }


def cpu_full_configs(threads):
    """The plain-method (64 x 128^2, fft, lags 1..63) and diff-mode (300 x 256^2
    uint8, 50 log-spaced lags) workloads run in full by the C restatement."""
    from fastddm_amd.lags import logspace_int
    from fastddm_amd.speckle import CONFIGS, ar1_speckle
    from oracle import port

    res = {}
    img = ar1_speckle(**CONFIGS["cpu_128"])
    L = img.shape[0]
    t0 = time.perf_counter()
    port.ddm_fft(img, np.arange(1, L), 128, 128, 128, None, chunk_size=65 * 128, threads=threads)
    dt = time.perf_counter() - t0
    res["fft_64x128x128"] = {"s": dt, "frames_per_s": L / dt}
    img = ar1_speckle(**CONFIGS["diff_256"])
    L = img.shape[0]
    t0 = time.perf_counter()
    port.ddm_diff(img, logspace_int(L - 1, num=50, endpoint=True), 256, 256, None, threads=threads)
    dt = time.perf_counter() - t0
    res["diff_300x256x256"] = {"s": dt, "frames_per_s": L / dt}
    return res


def cpu_baselines(img, lags, nx, ny, nt, win, L, nq):
    """cpu_baseline of the bench line: the 1-thread restatement (the reference
    cpp core is single-threaded) plus the all-core run and the two small
    workloads in full, with the host CPU named."""
    from oracle import port

    if not os.path.isfile(port._LIB_PATH):
        port.build()
    allc = max(1, min(16, len(os.sched_getaffinity(0))))
    base = cpu_baseline(img, lags, nx, ny, nt, win, L, nq, threads=1)
    base["cpu_model"] = _cpu_model()
    base["all_cores"] = cpu_baseline(img, lags, nx, ny, nt, win, L, nq, threads=allc)
    base["full_workloads"] = {"threads_1": cpu_full_configs(1), f"threads_{allc}": cpu_full_configs(allc)}
    return base


def batch_main(args):
    """Uneven stacks in one launch: 48 stacks, (H, W) cycling through 96x96,
    150x200, 384x384, 480x640, frames log-uniform in [50, 1500] (seed 4),
    fft mode, all lags, fp32c, float64 output; uint16 inputs made on the
    device (uniform integers in [800, 1200)).  One step = one replay of the
    batch graph; the same plans run one after another without the graph are
    timed beside it."""
    from fastddm_amd import _core_hip as core
    from fastddm_amd import next_fast_len

    core.set_device(0)
    rng = np.random.default_rng(4)
    shapes = [(96, 96), (150, 200), (384, 384), (480, 640)]
    n = 48
    Ls = np.exp(rng.uniform(np.log(50), np.log(1500), n)).round().astype(int)
    items = []
    t0 = time.perf_counter()
    for k in range(n):
        H, W = shapes[k % 4]
        L = int(Ls[k])
        img = core.DeviceBuffer(L * H * W * 2)
        core.fill_uniform(img, L * H * W, np.uint16, 7000 + k, 800, 1200)
        lags = np.arange(1, L)
        nt = next_fast_len(2 * L, "hip")
        plan = core.Plan("fft", np.uint16, (L, H, W), lags, W, H, nt, None, "fp32c")
        out = core.DeviceBuffer(plan.output_size())
        plan.destroy()
        items.append(dict(mode="fft", img=img, shape=(L, H, W), dtype=np.uint16, lags=lags, nx=W, ny=H,
                          nt=nt, precision="fp32c", out=out))
    batch = core.Batch(items, autotune=True)
    log(f"batch of {n} stacks built (plans, autotune, graph) in {time.perf_counter() - t0:.1f} s")
    stream = core.Stream()

    def timed(fn):
        for _ in range(args.warmup):
            fn(stream)
        stream.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            fn(stream)
        stream.synchronize()
        return (time.perf_counter() - t) / args.steps * 1e3

    ms_graph = timed(batch.launch)
    ms_eager = timed(batch.launch_eager)
    tbs = batch.frames_per_block()
    kinds = batch.spatial_kernels()
    frames = int(Ls.sum())
    batch.close()
    stream.destroy()
    for it in items:
        it["img"].free()
        it["out"].free()
    return {
        "metric": "batched fft-mode frames/s, 48 uneven stacks in one graph launch",
        "value": frames / (ms_graph / 1e3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_graph,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "device-made uniform uint16 stacks (seeded hash), shapes and lengths seeded",
        "config": {"workload": ("48 stacks, (H, W) cycling 96x96 / 150x200 / 384x384 / 480x640, "
                                "frames log-uniform in [50, 1500], fft mode, all lags, fp32c, float64 out"),
                   "frames": frames, "stacks": n, "parallelism": "single GPU"},
        "eager_ms_per_step": ms_eager,
        "graph_speedup_vs_eager": ms_eager / ms_graph,
        "frames_per_block_by_shape": {f"{h}x{w}": tbs[k] for k, (h, w) in enumerate(shapes)},
        "mixed_radix_rows_cols_by_shape": {f"{h}x{w}": list(kinds[k]) for k, (h, w) in enumerate(shapes)},
    }


def precision_main(args):
    """Error of D, PS and VAR against the FP64 oracle at the default workload
    (2000 x 512^2 uint16 speckle, lags 1..1999, Blackman-Harris, nt = 4000)
    for the fp64, fp32 and fp32c GPU paths: per lag, the median and max over q
    of |x - x_ref| / |x_ref| (x_ref != 0), overall, and at q = 0.  x_ref is the
    FP64 C restatement of the reference cpp core (oracle/ddm_port.c, OpenMP)."""
    from fastddm_amd import _core_hip as core
    from fastddm_amd.speckle import ar1_speckle
    from fastddm_amd.window import blackman_harris
    from oracle import port

    if not os.path.isfile(port._LIB_PATH):
        port.build()
    L, S = 2000, 512
    img = ar1_speckle(L, S, S, sigma=1.5, diffusivity=0.2, mean=1000.0, std=150.0, dtype=np.uint16, seed=2)
    win = blackman_harris((S, S))
    lags = np.arange(1, L)
    threads = min(16, len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    ref = port.ddm_fft(img, lags, S, S, 2 * L, win, chunk_size=257, threads=threads)
    log(f"FP64 oracle ({threads} threads): {time.perf_counter() - t0:.1f} s")
    core.set_device(0)
    q0 = (S // 2) * (S // 2 + 1)  # q = 0 sits at row ny/2 after the fftshift
    result = {"workload": "2000 x 512^2 uint16 speckle, lags 1..1999, Blackman-Harris, nt=4000, float64 out",
              "reference": f"oracle/ddm_port.c FP64, {threads} threads", "paths": {}}
    per_lag = {}
    for prec in ("fp64", "fp32", "fp32c"):
        out = core.ddm_fft(img, lags, S, S, 4000, win, precision=prec)
        a, r = out.reshape(out.shape[0], -1), ref.reshape(ref.shape[0], -1)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.abs(a - r) / np.abs(r)
        rel[r == 0] = np.nan
        med = np.nanmedian(rel, axis=1)
        mx = np.nanmax(rel, axis=1)
        per_lag[prec] = (med, mx)
        d = rel[:-2]
        result["paths"][prec] = {
            "D_median": float(np.nanmedian(d)), "D_max": float(np.nanmax(d)),
            "D_max_at_lag": int(lags[int(np.nanargmax(np.nanmax(d, axis=1)))]),
            "PS_median": float(med[-2]), "PS_max": float(mx[-2]),
            "VAR_median": float(med[-1]), "VAR_max": float(mx[-1]),
            "q0_D_max": float(np.nanmax(d[:, q0])), "q0_D_median": float(np.nanmedian(d[:, q0])),
            "per_lag_median_max": float(np.max(med[:-2])),
        }
        log(f"{prec}: {result['paths'][prec]}")
        del out, a, rel, d
    pick = [1, 2, 3, 5, 10, 20, 50, 100, 200, 500, 1000, 1500, 1900, 1968, 1990, 1999]
    result["per_lag"] = {str(l): {p: [float(per_lag[p][0][l - 1]), float(per_lag[p][1][l - 1])] for p in per_lag}
                         for l in pick}
    return result


def serve_main(args):
    """Serving run: Poisson arrivals of mixed analyses into the request server.

    500 requests at a fixed rate; size uniform in {128, 256, 512, 1024}^2,
    frames log-uniform in [100, 4000], 80 % fft (all lags) / 20 % diff (50
    log-spaced lags), fp32c, uint16, float64 output kept on device until the
    client releases it; five requests are forced to 1024^2 x 4000 fft.  Inputs
    are device-resident stacks of 4000 frames per size (uniform integers in
    [800, 1200) made on the device); request k uses the first L_k frames.
    Latency = completion seen by the server - submission (plan creation
    included).
    """
    import threading

    from fastddm_amd import _core_hip as core
    from fastddm_amd import next_fast_len
    from fastddm_amd.lags import logspace_int

    core.set_device(0)
    rng = np.random.default_rng(5)
    n, rate = args.serve_requests, args.serve_rate
    sizes = np.array([128, 256, 512, 1024])
    size = sizes[rng.integers(0, 4, n)]
    L = np.exp(rng.uniform(np.log(100), np.log(4000), n)).round().astype(int)
    mode = np.where(rng.random(n) < 0.8, "fft", "diff")
    forced = np.linspace(0, n - 1, 7).astype(int)[1:-1]  # five evenly spread
    size[forced], L[forced], mode[forced] = 1024, 4000, "fft"
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))

    stacks = {}
    host = args.serve_host_inputs
    for k, sz in enumerate(sizes):
        sz = int(sz)
        if host:
            # host stacks: 64 random frames repeated (the values do not matter
            # for latency); each request's frames cross the host link at admission
            base = np.random.default_rng(1000 + k).integers(800, 1200, (64, sz, sz)).astype(np.uint16)
            stacks[sz] = np.ascontiguousarray(np.tile(base, (4000 // 64 + 1, 1, 1))[:4000])
        else:
            b = core.DeviceBuffer(4000 * sz * sz * 2)
            core.fill_uniform(b, 4000 * sz * sz, np.uint16, 1000 + k, 800, 1200)
            stacks[sz] = b
    srv = core.Server(max_slots=args.serve_slots)
    rid_of = {}
    done = {}
    lock = threading.Lock()

    def drain():
        while True:
            with lock:
                if len(done) == n:
                    return
            rid = srv.next_done(0.5)
            if rid == 0:
                continue
            st, t = srv.status(rid)
            with lock:
                done[rid] = (st, t)
            srv.release(rid)

    th = threading.Thread(target=drain, daemon=True)
    th.start()
    t0 = time.perf_counter()
    for i in range(n):
        wait = arrivals[i] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        sz, l = int(size[i]), int(L[i])
        lags = np.arange(1, l) if mode[i] == "fft" else logspace_int(l - 1, num=50, endpoint=True)
        nt = next_fast_len(2 * l, "hip") if mode[i] == "fft" else 0
        if host:
            rid = srv.submit(mode[i], stacks[sz][:l], lags, sz, sz, nt, precision="fp32c")
        else:
            rid = srv.submit(mode[i], stacks[sz], lags, sz, sz, nt, precision="fp32c", img_on_device=True,
                             shape=(l, sz, sz), dtype=np.uint16)
        rid_of[i] = rid
    th.join(timeout=600)
    stats = srv.stats()
    srv.close()
    if not host:
        for b in stacks.values():
            b.free()
    if len(done) != n:
        raise RuntimeError(f"only {len(done)} of {n} requests finished")
    lat = np.empty(n)
    first, last = np.inf, 0.0
    for i in range(n):
        st, (ts, ta, td) = done[rid_of[i]]
        if st != core.Server.DONE:
            raise RuntimeError(f"request {i} failed")
        lat[i] = (td - ts) * 1e3
        first, last = min(first, ts), max(last, td)
    frames = int(L.sum())
    per_size = {}
    for sz in sizes:
        m = size == sz
        per_size[f"{sz}x{sz}"] = {"requests": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50)),
                                 "p99_ms": float(np.percentile(lat[m], 99))}
    return {
        "requests": n,
        "arrival_rate_per_s": rate,
        "slots": args.serve_slots,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "sustained_frames_per_s": frames / (last - first),
        "frames": frames,
        "span_s": last - first,
        "fft_fraction": float(np.mean(mode == "fft")),
        "forced_1024x4000": int(len(forced)),
        "arena_GB": stats["arena_bytes"] / 1e9,
        "arena_peak_GB": stats["arena_peak"] / 1e9,
        "plans": stats["plans"],
        "per_size": per_size,
        "inputs": "host" if host else "device",
        "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES", "runtime default"),
        "note": (("host uint16 stacks copied to the device at admission (inside the latency)" if host else
                  "device-resident uint16 stacks (uniform integers)") +
                 ", fp32c, float64 results kept on device until released; latency includes plan creation; "
                 "requests above 512 MB of spectra run as launches of 256 MB each"),
    }


def _subset_rows(nl, L, lags):
    """Output rows compared in the large leg: every 16th lag, every lag with
    at most 64 frame pairs, PS and VAR."""
    lags = np.asarray(lags)
    rows = set(range(0, nl, 16)) | set(np.nonzero(L - lags <= 64)[0].tolist()) | {nl, nl + 1}
    return np.array(sorted(rows))


def large_leg(core, args):
    """The large workload: 4000 frames of 1024x1024 uint16 AR(1) speckle
    (sigma 1.5, D 0.2, mean 1000, std 150, seed 6), lags 1..3999,
    Blackman-Harris, nt 8000, float64 output, input resident in HBM; fp32c
    and fp64 timed with per-stage HIP events, fp32c error against fp64 on a
    subset of output rows."""
    from fastddm_amd.speckle import ar1_speckle
    from fastddm_amd.window import blackman_harris

    L, S = 4000, 1024
    t0 = time.perf_counter()
    img = ar1_speckle(L, S, S, sigma=1.5, diffusivity=0.2, mean=1000.0, std=150.0, dtype=np.uint16, seed=6)
    log(f"large leg: generated {img.shape} speckle in {time.perf_counter() - t0:.1f} s")
    win = blackman_harris((S, S))
    lags = np.arange(1, L)
    nt = 2 * L
    nxh = S // 2 + 1
    nq = S * nxh
    d_img = core.DeviceBuffer.from_array(img)
    del img
    rows = _subset_rows(len(lags), L, lags)
    res = {"workload": (f"fft mode, {L} frames of {S}x{S} uint16 speckle, lags 1..{L - 1}, Blackman-Harris, "
                        f"nt={nt} (time transform length {_time_transform_length(L)}), float64 output"),
           "frames": L, "height": S, "width": S}
    outs = {}
    for prec, steps in (("fp32c", args.large_steps), ("fp64", max(2, args.large_steps // 2))):
        plan = core.Plan("fft", np.uint16, (L, S, S), lags, S, S, nt, win, prec)
        d_out = core.DeviceBuffer(plan.output_size())
        d_ws = core.DeviceBuffer(plan.workspace_size())
        stream = core.Stream()
        plan.execute(d_img, d_out, d_ws, stream)
        stream.synchronize()
        plan.timing_begin(steps)
        t = time.perf_counter()
        for _ in range(steps):
            plan.execute(d_img, d_out, d_ws, stream)
        stream.synchronize()
        ms = (time.perf_counter() - t) / steps * 1e3
        st, n = plan.timing_end()
        st = {k: v / max(n, 1) for k, v in st.items()}
        row_bytes = nq * 8
        outs[prec] = np.stack([d_out.read(int(r) * row_bytes, (nq,), np.float64) for r in rows])
        c = 8 if prec == "fp32c" else 16
        b = algorithmic_bytes(L, S, S, 2, nq, len(lags), c, 8, True)
        res[prec] = {"ms_per_step": ms, "frames_per_s": L / (ms / 1e3), "stages_ms": st,
                     "stage_hbm": stage_fractions(st, b),
                     "whole_path_frac": b["whole"] / (ms / 1e3) / 1e9 / HBM_PEAK_GBS}
        log(f"large leg {prec}: {ms:.2f} ms/step, stages {st}")
        for buf in (d_out, d_ws):
            buf.free()
        stream.destroy()
        plan.destroy()
    d_img.free()
    a, r = outs["fp32c"], outs["fp64"]
    m = r != 0
    rel = np.abs(a[m] - r[m]) / np.abs(r[m])
    res["speedup_fp32c_vs_fp64"] = res["fp64"]["ms_per_step"] / res["fp32c"]["ms_per_step"]
    res["fp32c_vs_fp64_rel_err"] = {"median": float(np.median(rel)), "max": float(rel.max()),
                                   "rows_compared": int(len(rows)),
                                   "rows": "every 16th lag, every lag with <= 64 frame pairs, PS, VAR"}
    return res


def diff_leg(core, args):
    """Diff mode: 300 frames of 256x256 uint8 speckle (sigma 2, D 1, mean 120,
    std 25, clipped 0..255, seed 3), 50 log-spaced lags in 1..299, no window,
    fp32c and fp64, input resident in HBM."""
    from fastddm_amd.lags import logspace_int
    from fastddm_amd.speckle import CONFIGS, ar1_speckle

    img = ar1_speckle(**CONFIGS["diff_256"])
    L, H, W = img.shape
    lags = logspace_int(L - 1, num=50, endpoint=True)
    nq = H * (W // 2 + 1)
    res = {"workload": f"diff mode, {L} frames of {H}x{W} uint8 speckle, 50 log-spaced lags in 1..{L - 1}, no window"}
    outs = {}
    for prec in ("fp32c", "fp64"):
        ms, st, out = run_gpu(core, img, lags, W, H, 0, None, prec, args.steps, args.warmup, keep_output=True,
                              mode="diff")
        outs[prec] = out
        pairs = nq * int(np.sum(L - lags))
        tflops = 6.0 * pairs / (st["time_or_diff"] / 1e3) / 1e12
        c = 8 if prec == "fp32c" else 16
        b = algorithmic_bytes(L, H, W, 1, nq, len(lags), c, 8, False)
        res[prec] = {"ms_per_step": ms, "frames_per_s": L / (ms / 1e3), "stages_ms": st,
                     "stage_hbm": stage_fractions(st, b),
                     "valu": {"kernel": "diff", "achieved": tflops, "peak": FP64_VECTOR_PEAK_TFLOPS,
                              "unit": "TFLOP/s", "frac": tflops / FP64_VECTOR_PEAK_TFLOPS, "differences": pairs}}
        log(f"diff leg {prec}: {ms:.3f} ms/step, stages {st}")
    m = outs["fp64"] != 0
    rel = np.abs(outs["fp32c"][m] - outs["fp64"][m]) / np.abs(outs["fp64"][m])
    res["fp32c_vs_fp64_rel_err"] = {"median": float(np.median(rel)), "max": float(rel.max())}
    return res


def _traffic_entry(dom, L, S):
    """HBM traffic per launch of the dominant kernel, from the stored rocprofv3
    counter run (not measured in this run; the entry names its source)."""
    tf = os.path.join(ROOT, "profiles", "traffic.json")
    if not (os.path.isfile(tf) and dom == "time_wk" and (L, S) == (2000, 512)):
        return None, None
    with open(tf) as f:
        t = json.load(f)
    entry = t.get("k_time_fp32c_512x2000")
    if not entry:
        return None, None
    return entry.get("bytes_per_launch"), entry.get("source")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", choices=["fft", "diff", "batch"], default="fft",
                    help="fft: 2000 x 512^2 uint16, all lags (default); diff: 300 x 256^2 uint8, "
                         "50 log-spaced lags, no window")
    ap.add_argument("--frames", type=int, default=None)
    ap.add_argument("--size", type=int, default=None)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-serve", action="store_true")
    ap.add_argument("--skip-large", action="store_true", help="skip the 1024^2 x 4000 and 48-stack batch legs")
    ap.add_argument("--skip-diff-leg", action="store_true", help="skip the diff-mode leg")
    ap.add_argument("--skip-host", action="store_true", help="skip the host-to-host and streamed-ingest legs")
    ap.add_argument("--skip-batch", action="store_true", help="skip the 48-stack batch leg")
    ap.add_argument("--large-steps", type=int, default=5)
    ap.add_argument("--serve-only", action="store_true", help="run only the serving measurement")
    ap.add_argument("--precision-report", action="store_true",
                    help="only report fp64 / fp32 / fp32c errors against the FP64 oracle")
    ap.add_argument("--serve-requests", type=int, default=500)
    ap.add_argument("--serve-rate", type=float, default=100.0, help="Poisson arrival rate (1/s)")
    ap.add_argument("--serve-slots", type=int, default=8)
    ap.add_argument("--serve-host-inputs", action="store_true",
                    help="serve from host stacks (the H2D of each request inside its latency)")
    ap.add_argument("--serve-hw-queues", type=int, default=0,
                    help="GPU_MAX_HW_QUEUES for the serving process (0: one per slot, at most 32; "
                         "-1: the runtime default)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run 3 device-to-device copies of 1 GiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(whole stacks would be routed to GPUs independently). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    if args.serve_only:
        print(json.dumps(serve_main(args)), flush=True)
        return
    if args.workload == "batch":
        print(json.dumps(batch_main(args)), flush=True)
        return
    if args.precision_report:
        print(json.dumps(precision_main(args)), flush=True)
        return
    serve = None
    if not args.skip_serve and not args.copy_calibration and args.workload == "fft":
        # in a child process started before this one touches the GPU
        import subprocess

        cmd = [sys.executable, os.path.abspath(__file__), "--serve-only",
               "--serve-requests", str(args.serve_requests), "--serve-rate", str(args.serve_rate),
               "--serve-slots", str(args.serve_slots)]
        if args.serve_host_inputs:
            cmd.append("--serve-host-inputs")
        env = dict(os.environ)
        if 0 < args.serve_hw_queues <= 32:
            env["GPU_MAX_HW_QUEUES"] = str(args.serve_hw_queues)
        elif args.serve_hw_queues == 0:
            # one hardware queue per worker stream (the runtime's default is 4)
            env["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.serve_slots)))
        t0 = time.perf_counter()
        try:
            res = subprocess.run(cmd, capture_output=True, text=True, timeout=420, env=env)
            lines = [l for l in res.stdout.splitlines() if l.startswith("{")]
            serve = json.loads(lines[-1]) if res.returncode == 0 and lines else {
                "error": f"serving run exited with {res.returncode}: {res.stderr[-400:]}"}
        except subprocess.TimeoutExpired:
            serve = {"error": "serving run exceeded 420 s"}
        log(f"serving run: {time.perf_counter() - t0:.1f} s")

    from fastddm_amd import _core_hip as core
    from fastddm_amd import next_fast_len
    from fastddm_amd.speckle import ar1_speckle
    from fastddm_amd.lags import logspace_int
    from fastddm_amd.window import blackman_harris

    if core.get_num_devices() < 1:
        raise RuntimeError("no AMD GPU visible")
    core.set_device(0)
    if args.copy_calibration:
        n = 1 << 30
        a, b = core.DeviceBuffer(n), core.DeviceBuffer(n)
        for _ in range(3):
            b.copy_from(a, n)
        log("copied 3 x 1 GiB device to device")
        return

    diff = args.workload == "diff"
    L = args.frames or (300 if diff else 2000)
    S = args.size or (256 if diff else 512)
    t0 = time.perf_counter()
    if diff:
        img = ar1_speckle(L, S, S, sigma=2.0, diffusivity=1.0, mean=120.0, std=25.0,
                          dtype=np.uint8, seed=3, clip=(0, 255))
        win = None
        lags = logspace_int(L - 1, num=50, endpoint=True)
    else:
        img = ar1_speckle(L, S, S, sigma=1.5, diffusivity=0.2, mean=1000.0, std=150.0,
                          dtype=np.uint16, seed=2)
        win = blackman_harris((S, S))
        lags = np.arange(1, L)
    log(f"generated {img.shape} {img.dtype} speckle in {time.perf_counter() - t0:.1f} s")
    nx = next_fast_len(S, "hip")
    ny = nx
    nt = next_fast_len(2 * L, "hip")
    nxh = nx // 2 + 1
    nq = ny * nxh

    mode = "diff" if diff else "fft"
    ms32, st32, out32 = run_gpu(core, img, lags, nx, ny, nt, win, "fp32c", args.steps, args.warmup,
                                keep_output=not args.skip_fp64, mode=mode)
    log(f"fp32c: {ms32:.3f} ms/step, stages {st32}")
    # host numpy in -> host numpy out through the C ABI (PCIe both ways, device
    # buffers cached after the first call); reported beside, never as value
    e2e_s, ingest_s, ingest_parts = None, [], []
    if not args.skip_host:
        e2e = []
        for _ in range(3):
            t0 = time.perf_counter()
            if diff:
                core.ddm_diff(img, lags, nx, ny, win, precision="fp32c")
            else:
                core.ddm_fft(img, lags, nx, ny, nt, win, precision="fp32c")
            e2e.append(time.perf_counter() - t0)
        e2e_s = min(e2e[1:])
        core.release_cache()
        log(f"host-to-host fp32c: {e2e_s * 1e3:.1f} ms per call")
        # the same analysis from host frames pushed in batches of 64 (streamed
        # ingest: pinned chunks, DMA and row/column passes overlapped)
        for _ in range(2):
            t0 = time.perf_counter()
            ing = core.Ingest(mode, img.dtype, img.shape, lags, nx, ny, 0 if diff else nt, win, "fp32c")
            t1 = time.perf_counter()
            for k in range(0, L, 64):
                ing.push(img[k:k + 64])
            t2 = time.perf_counter()
            ing.finish()
            t3 = time.perf_counter()
            ing.close()
            ingest_s.append(time.perf_counter() - t0)
            ingest_parts.append({"begin_ms": (t1 - t0) * 1e3, "push_ms": (t2 - t1) * 1e3, "finish_ms": (t3 - t2) * 1e3,
                                 "close_ms": (ingest_s[-1] - (t3 - t0)) * 1e3})
        log(f"streamed ingest fp32c: {min(ingest_s) * 1e3:.1f} ms per analysis")
        # the same with a result array whose pages are already mapped (the
        # difference is the first-touch cost of a fresh 2 GB host array)
        ing = core.Ingest(mode, img.dtype, img.shape, lags, nx, ny, 0 if diff else nt, win, "fp32c")
        dst = np.zeros(ing.out_shape, dtype=ing.out_dtype)
        t0 = time.perf_counter()
        for k in range(0, L, 64):
            ing.push(img[k:k + 64])
        t1 = time.perf_counter()
        ing.finish(out=dst)
        ingest_warm = {"push_ms": (t1 - t0) * 1e3, "finish_ms": (time.perf_counter() - t1) * 1e3}
        ing.close()
        del dst
        log(f"streamed ingest into a mapped result: {ingest_warm}")
    fp64 = None
    b64 = algorithmic_bytes(L, S, S, img.itemsize, nq, len(lags), 16, 8, win is not None)
    if not args.skip_fp64:
        ms64, st64, out64 = run_gpu(core, img, lags, nx, ny, nt, win, "fp64", args.steps,
                                    args.warmup, keep_output=True, mode=mode)
        log(f"fp64: {ms64:.3f} ms/step, stages {st64}")
        d32, d64 = out32[:-2].ravel(), out64[:-2].ravel()
        m = d64 != 0
        rel = np.abs(d32[m] - d64[m]) / np.abs(d64[m])
        fp64 = {
            "frames_per_s": L / (ms64 / 1e3),
            "ms_per_step": ms64,
            "stages_ms": st64,
            "stage_hbm": stage_fractions(st64, b64),
            "fp32c_vs_fp64_gpu_rel_err": {"median": float(np.median(rel)), "max": float(rel.max())},
        }
        del out32, out64, d32, d64, rel

    b = algorithmic_bytes(L, S, S, img.itemsize, nq, len(lags), 8, 8, win is not None)
    kernels = {
        ("diff" if diff else "time_wk"): (st32["time_or_diff"], b["time"]),
        "rows": (st32["row_pass"], b["rows"]),
        "columns": (st32["column_pass"], b["cols"]),
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    dom_ms, dom_bytes = kernels[dom]
    achieved = dom_bytes / (dom_ms / 1e3) / 1e9
    traffic, traffic_src = _traffic_entry(dom, L, S)
    tlen = None if diff else _time_transform_length(L)

    result = {
        "metric": "fft-mode frames/s + HBM roofline %, 512^2x2000 fp32 vs fp64; serve p50/p99",
        "value": L / (ms32 / 1e3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms32,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": (f"synthetic AR(1) Brownian speckle, {img.dtype}, seed {3 if diff else 2} "
                 "(device-resident input)"),
        "config": {
            "workload": (f"diff mode, {L} frames of {S}x{S} uint8, {len(lags)} log-spaced lags in "
                         f"1..{L - 1}, no window, precision fp32c, float64 output" if diff else
                         f"fft mode, {L} frames of {S}x{S} uint16, lags 1..{L - 1}, Blackman-Harris "
                         f"window, nt={nt} (time transform run at length {tlen}: D is the same for "
                         f"any zero-padded length >= 2L-1), precision fp32c, float64 output"),
            "frames": L, "height": S, "width": S, "nt": nt, "time_transform_length": tlen,
            "lags": len(lags), "precision": "fp32c", "parallelism": "single GPU",
        },
        "stages_ms": st32,
        "stage_hbm": stage_fractions(st32, b),
        "roofline": {
            "kernel": dom,
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_src,
            "algorithmic_bytes": dom_bytes,
        },
        "host_to_host": None if e2e_s is None else {
            "frames_per_s": L / e2e_s,
            "ms_per_call": e2e_s * 1e3,
            "note": (f"{'ddm_diff' if diff else 'ddm_fft'} through the C ABI from a host numpy stack to a "
                     f"host numpy result (H2D of the {img.dtype} frames and D2H of the float64 output "
                     "included)"),
        },
        "ingest": None if not ingest_s else {
            "frames_per_s": L / min(ingest_s),
            "ms_per_analysis": min(ingest_s) * 1e3,
            "phases_ms": {k: round(v, 2) for k, v in ingest_parts[int(np.argmin(ingest_s))].items()},
            "into_mapped_result_ms": {k: round(v, 2) for k, v in ingest_warm.items()},
            "note": ("host frames pushed 64 at a time into the streamed ingest (device buffers allocated "
                     "per analysis, included), result copied back to a host array"),
        },
        "whole_path": {
            "algorithmic_bytes": b["whole"],
            "effective_GBs": b["whole"] / (ms32 / 1e3) / 1e9,
            "frac_of_hbm_peak": b["whole"] / (ms32 / 1e3) / 1e9 / HBM_PEAK_GBS,
        },
    }
    if diff:
        # the diff kernel is compute bound: its FP64 rate beside the HBM line
        # (6 flops per difference: 2 subtractions, 2 squares, 2 additions)
        pairs = nq * int(np.sum(L - np.asarray(lags)))
        tflops = 6.0 * pairs / (st32["time_or_diff"] / 1e3) / 1e12
        result["roofline"]["valu"] = {"achieved": tflops, "peak": FP64_VECTOR_PEAK_TFLOPS, "unit": "TFLOP/s",
                                      "frac": tflops / FP64_VECTOR_PEAK_TFLOPS, "differences": pairs}
    if serve is not None:
        result["serve"] = serve
    if fp64 is not None:
        result["fp64"] = fp64
        result["speedup_vs_fp64"] = fp64["ms_per_step"] / ms32
    if not diff and not args.skip_large and args.frames is None and args.size is None:
        t0 = time.perf_counter()
        result["large_1024x4000"] = large_leg(core, args)
        log(f"large leg: {time.perf_counter() - t0:.1f} s")
    if not diff and not args.skip_diff_leg and args.frames is None and args.size is None:
        result["diff_300x256x256"] = diff_leg(core, args)
    if not diff and not args.skip_large and not args.skip_batch and args.frames is None and args.size is None:
        t0 = time.perf_counter()
        b = batch_main(argparse.Namespace(steps=5, warmup=2))
        result["batch_48_stacks"] = {
            "workload": b["config"]["workload"], "frames": b["config"]["frames"],
            "graph_ms": b["ms_per_step"], "eager_ms": b["eager_ms_per_step"], "frames_per_s": b["value"],
            "graph_speedup_vs_eager": b["graph_speedup_vs_eager"],
            "frames_per_block_by_shape": b["frames_per_block_by_shape"],
            "mixed_radix_rows_cols_by_shape": b["mixed_radix_rows_cols_by_shape"]}
        log(f"batch leg: {time.perf_counter() - t0:.1f} s")
    if not args.skip_cpu and not diff:
        t0 = time.perf_counter()
        result["cpu_baseline"] = cpu_baselines(img, lags, nx, ny, nt, win, L, nq)
        log(f"cpu baselines took {time.perf_counter() - t0:.1f} s")
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
