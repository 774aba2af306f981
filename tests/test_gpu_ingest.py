"""Streamed ingest (fddm_ingest_*): frames pushed in irregular pieces give the
result of the one-call host path bit for bit -- fft and diff mode, fp32c and
fp64, stacks whose length is not a multiple of the frame block, stacks of at
most 16 frames (the FP64 edge child then takes every frame), a result larger
than the pinned chunk buffers it is copied out through -- and
ddm_files on a multipage TIFF equals ddm on the array it holds."""

import numpy as np
import pytest
from PIL import Image

from fastddm_amd import next_fast_len
from fastddm_amd.speckle import ar1_speckle
from fastddm_amd.window import blackman_harris

pytestmark = pytest.mark.gpu

CASES = [
    # mode, precision, (L, H, W), window, pieces
    ("fft", "fp32c", (301, 64, 96), True, [1, 7, 33, 100]),
    ("fft", "fp64", (130, 48, 40), False, [64, 64]),
    ("diff", "fp32c", (90, 40, 40), True, [5]),
    ("fft", "fp32c", (12, 32, 32), False, [3, 4]),
    # 26 MB chunk buffers, 53 MB result: the result leaves in three staged pieces
    ("fft", "fp32c", (200, 256, 256), True, [64]),
]


@pytest.mark.parametrize("mode,precision,shape,window,pieces", CASES)
def test_ingest_equals_host_path(hip, mode, precision, shape, window, pieces):
    L, H, W = shape
    img = ar1_speckle(L, H, W, sigma=1.3, diffusivity=0.3, mean=1200.0, std=180.0, seed=L + W)
    win = blackman_harris((H, W)) if window else None
    lags = np.arange(1, L) if mode == "fft" else np.unique(np.geomspace(1, L - 1, 25).astype(int))
    nt = next_fast_len(2 * L, "hip")
    print(f"ingest {mode} {precision} {shape} pieces {pieces}", flush=True)
    if mode == "fft":
        ref = hip.ddm_fft(img, lags, W, H, nt, win, precision=precision)
    else:
        ref = hip.ddm_diff(img, lags, W, H, win, precision=precision)
    ing = hip.Ingest(mode, img.dtype, shape, lags, W, H, nt if mode == "fft" else 0, win, precision)
    t, k = 0, 0
    while t < L:
        n = min(pieces[k % len(pieces)], L - t)
        ing.push(img[t:t + n])
        t += n
        k += 1
    got = ing.finish()
    ing.close()
    assert np.array_equal(got, ref), np.max(np.abs(got - ref))


def test_ingest_refuses_wrong_counts(hip):
    img = np.zeros((20, 16, 16), np.uint16)
    ing = hip.Ingest("fft", np.uint16, img.shape, [1, 2], 16, 16, 40, None, "fp64")
    ing.push(img[:10])
    with pytest.raises(RuntimeError, match="of 20 frames"):
        ing.finish()
    with pytest.raises(RuntimeError, match="pushed"):
        ing.push(np.zeros((11, 16, 16), np.uint16))
    with pytest.raises(RuntimeError, match="frames of shape"):
        ing.push(np.zeros((1, 8, 16), np.uint16))
    ing.close()


def test_ddm_files_equals_ddm_on_array(hip, confocal, tmp_path):
    import fastddm_amd as fddm
    from fastddm_amd import utils

    path = str(tmp_path / "confocal.tif")
    ims = [Image.fromarray(f) for f in confocal]
    ims[0].save(path, save_all=True, append_images=ims[1:])
    lags = np.arange(1, 10)
    win = blackman_harris(confocal.shape[1:])
    for mode in ("fft", "diff"):
        ref = fddm.ddm(confocal, lags, core="hip", mode=mode, window=win, precision="fp32c")
        got = utils.ddm_files(path, lags, mode=mode, window=win, precision="fp32c", batch=6)
        assert np.array_equal(got._data, ref._data)
        assert np.array_equal(got.kx, ref.kx) and np.array_equal(got.ky, ref.ky)
        assert np.array_equal(got.tau, ref.tau)


def test_ingest_into_given_array(hip):
    """finish(out=...) writes the result into the caller's array (and refuses
    one of the wrong shape or dtype before touching the device)."""
    img = ar1_speckle(40, 32, 32, seed=9)
    lags = np.arange(1, 40)
    ref = hip.ddm_fft(img, lags, 32, 32, 80, None, precision="fp32c")
    ing = hip.Ingest("fft", img.dtype, img.shape, lags, 32, 32, 80, None, "fp32c")
    ing.push(img)
    with pytest.raises(RuntimeError, match="output array"):
        ing.finish(out=np.zeros((3, 32, 17)))
    dst = np.full(ing.out_shape, np.nan)
    got = ing.finish(out=dst)
    ing.close()
    assert got is dst and np.array_equal(dst, ref)
