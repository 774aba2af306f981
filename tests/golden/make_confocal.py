"""Writes tests/golden/confocal_20.npz: the first 20 frames of the reference's
test stack tests/test-imgs/confocal/0000.tif ... (256x256, 8-bit grayscale),
the input of the reference's cross-core tests (tests/test_ddm.py:14-16).

Run in a container that holds the reference checkout:
    python tests/golden/make_confocal.py /root/reference
"""
import os
import sys

import numpy as np
from PIL import Image

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = os.path.join(ref, "tests", "test-imgs", "confocal")
names = sorted(f for f in os.listdir(src) if f.endswith(".tif"))[:20]
frames = np.stack([np.asarray(Image.open(os.path.join(src, n))) for n in names])
assert frames.shape == (20, 256, 256) and frames.dtype == np.uint8, (frames.shape, frames.dtype)
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "confocal_20.npz")
np.savez_compressed(out, frames=frames, names=np.array(names))
print(out, frames.shape, frames.dtype, int(frames.min()), int(frames.max()))
