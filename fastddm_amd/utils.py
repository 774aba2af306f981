"""Image readers and the streamed analysis of image files.

The readers mirror the reference's (src/python/utils.py:23-257) for
grayscale stacks in (t, y, x) order, built on Pillow instead of tifffile /
scikit-image (not part of this package's dependencies):

* ``tiff2numpy(src, seq)``   -- a multipage TIFF, optionally a subset of pages
  (reference :23-118);
* ``images2numpy(fnames)``   -- one image per file (reference :121-174);
* ``read_images(src, seq)``  -- either of the above (reference :208-256).

``ddm_files`` analyses image files without holding the whole stack first:
pages are decoded one by one and pushed in batches into the streamed ingest
of the HIP core (``fddm_ingest_*``), whose DMA and row / column transforms run
while the next batch is decoded.  The result is the ``ImageStructureFunction``
that ``ddm(read_images(src), lags, core='hip', ...)`` returns, value for
value.
"""

from typing import Iterable, Iterator, List, Optional, Sequence, Union

import numpy as np

_MODES = {"L": np.uint8, "I;16": np.uint16, "I;16L": np.uint16, "I;16B": np.uint16, "I": np.int32, "F": np.float32}


def _page_array(img) -> np.ndarray:
    if img.mode not in _MODES:
        raise RuntimeError(f"Unsupported image mode {img.mode!r} (grayscale images only).")
    a = np.asarray(img)
    want = np.dtype(_MODES[img.mode])
    if img.mode == "I;16B":
        a = a.astype(want)  # byte-swapped to native order
    return np.ascontiguousarray(a, dtype=want)


def iter_tiff_frames(src: str, seq: Optional[Sequence[int]] = None) -> Iterator[np.ndarray]:
    """Yield the pages of a (multipage) TIFF as 2D arrays, in order, or the
    pages listed in ``seq``."""
    from PIL import Image

    with Image.open(src) as im:
        n = getattr(im, "n_frames", 1)
        pages = range(n) if seq is None else seq
        for k in pages:
            if not 0 <= k < n:
                raise RuntimeError(f"Page {k} outside the {n} pages of {src}.")
            im.seek(k)
            yield _page_array(im)


def tiff2numpy(src: str, seq: Optional[Sequence[int]] = None) -> np.ndarray:
    """Read a multipage TIFF (t, y, x) into an array, optionally only the pages
    in ``seq`` (reference utils.py:23-118, grayscale)."""
    frames = list(iter_tiff_frames(src, seq))
    if not frames:
        raise RuntimeError(f"No pages read from {src}.")
    return np.stack(frames)


def images2numpy(fnames: Sequence[str]) -> np.ndarray:
    """Read one image per file into a (t, y, x) array (reference utils.py:121-174,
    grayscale)."""
    from PIL import Image

    frames = []
    for f in fnames:
        with Image.open(f) as im:
            frames.append(_page_array(im))
    return np.stack(frames)


def read_images(src: Union[str, List[str]], seq: Optional[Sequence[int]] = None) -> np.ndarray:
    """A single (multipage) image file or a list of files (reference utils.py:208-256)."""
    if isinstance(src, list):
        return images2numpy(src)
    if isinstance(src, str):
        if src.endswith(".nd2"):
            raise RuntimeError("ND2 files need nd2reader, which this package does not use.")
        return tiff2numpy(src, seq=seq)
    raise RuntimeError(f"Failed to open {src}.")


def _frames_of(src, seq) -> (Iterable[np.ndarray], int):
    from PIL import Image

    if isinstance(src, list):
        def files():
            for f in src:
                with Image.open(f) as im:
                    yield _page_array(im)

        return files(), len(src)
    if isinstance(src, str):
        with Image.open(src) as im:
            n = getattr(im, "n_frames", 1)
        return iter_tiff_frames(src, seq), (n if seq is None else len(seq))
    raise RuntimeError(f"Failed to open {src}.")


def ddm_files(src: Union[str, List[str]], lags, *, mode: str = "fft", window: Optional[np.ndarray] = None,
              seq: Optional[Sequence[int]] = None, batch: int = 64, gpu_id: int = 0, precision: str = "fp64"):
    """``ddm`` of image files, decoded page by page while the GPU transforms
    earlier pages (streamed ingest).  Same checks, padding and result as
    ``fastddm_amd.ddm(read_images(src, seq), lags, core='hip', mode=mode,
    window=window, precision=precision)``."""
    from . import _core_hip
    from ._config import DTYPE
    from ._ddm import _prepare_lags
    from ._fftopt import next_fast_len
    from .imagestructurefunction import ImageStructureFunction

    if mode not in ("fft", "diff"):
        raise RuntimeError(f"Unknown mode '{mode}' selected. Only possible options are ['diff', 'fft'].")
    frames, length = _frames_of(src, seq)
    it = iter(frames)
    first = next(it)
    dim_y, dim_x = first.shape
    lg = _prepare_lags(lags, length)
    if window is not None and np.size(window) > 0 and window.shape != (dim_y, dim_x):
        raise RuntimeError(
            f"Window with shape {window.shape} incompatible with image shape {(dim_y, dim_x)}."
        )
    nx, ny = next_fast_len(dim_x, "hip"), next_fast_len(dim_y, "hip")
    nt = next_fast_len(2 * length, "hip")
    win = None if window is None or np.size(window) == 0 else np.asarray(window).astype(DTYPE)
    _core_hip.set_device(gpu_id)
    ing = _core_hip.Ingest(mode, first.dtype, (length, dim_y, dim_x), lg, nx, ny, nt if mode == "fft" else 0, win,
                           precision=precision, out_dtype=DTYPE)
    try:
        buf = np.empty((batch, dim_y, dim_x), dtype=first.dtype)
        buf[0] = first
        k = 1
        for fr in it:
            if fr.shape != (dim_y, dim_x) or fr.dtype != first.dtype:
                raise RuntimeError("All images must have the same shape and type.")
            buf[k] = fr
            k += 1
            if k == batch:
                ing.push(buf)
                k = 0
        if k:
            ing.push(buf[:k])
        data = ing.finish()
    finally:
        ing.close()
    kx = 2 * np.pi * np.fft.fftfreq(nx)[: nx // 2 + 1]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny))
    return ImageStructureFunction(data, kx.astype(DTYPE), ky.astype(DTYPE), nx, ny, lg.astype(DTYPE))
