"""Image readers (fastddm_amd.utils, reference src/python/utils.py:23-257):
multipage TIFF and one-file-per-frame stacks written with Pillow and read
back exactly, page subsets, and the refused inputs (CPU only)."""

import numpy as np
import pytest
from PIL import Image

from fastddm_amd import utils


def _write_stack(path, frames):
    ims = [Image.fromarray(f) for f in frames]
    ims[0].save(path, save_all=True, append_images=ims[1:])


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16])
def test_multipage_tiff_round_trip(tmp_path, dtype):
    rng = np.random.default_rng(3)
    hi = 255 if dtype == np.uint8 else 60000
    frames = rng.integers(0, hi, (9, 17, 23)).astype(dtype)
    path = str(tmp_path / "stack.tif")
    _write_stack(path, frames)
    got = utils.tiff2numpy(path)
    assert got.dtype == dtype and np.array_equal(got, frames)
    assert np.array_equal(utils.read_images(path), frames)
    assert np.array_equal(utils.tiff2numpy(path, seq=range(2, 7)), frames[2:7])
    assert np.array_equal(np.stack(list(utils.iter_tiff_frames(path, seq=[8, 0]))), frames[[8, 0]])
    with pytest.raises(RuntimeError):
        utils.tiff2numpy(path, seq=[9])


def test_one_file_per_frame(tmp_path):
    rng = np.random.default_rng(4)
    frames = rng.integers(0, 4000, (5, 12, 10)).astype(np.uint16)
    names = []
    for k, f in enumerate(frames):
        names.append(str(tmp_path / f"img_{k:03d}.tif"))
        Image.fromarray(f).save(names[-1])
    assert np.array_equal(utils.images2numpy(names), frames)
    assert np.array_equal(utils.read_images(names), frames)


def test_refused_inputs(tmp_path):
    with pytest.raises(RuntimeError, match="ND2"):
        utils.read_images(str(tmp_path / "movie.nd2"))
    with pytest.raises(RuntimeError):
        utils.read_images(42)
    rgb = str(tmp_path / "rgb.tif")
    Image.fromarray(np.zeros((4, 4, 3), np.uint8)).save(rgb)
    with pytest.raises(RuntimeError, match="grayscale"):
        utils.tiff2numpy(rgb)
