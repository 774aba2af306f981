"""fastddm_amd: FastDDM's image structure function engine for AMD MI355X (gfx950).

The ``ddm()`` front end, its ``ImageStructureFunction`` result and the input
helpers (``window``, ``lags``) and the azimuthal average of the result mirror
somexlab/fastddm; the compute runs in hand-written HIP kernels behind
``core='hip'`` (``libfddm_hip.so``).
"""

from ._config import *  # noqa: F401,F403
from ._ddm import ddm, _backend  # noqa: F401
from ._fftopt import next_fast_len  # noqa: F401
from .imagestructurefunction import ImageStructureFunction  # noqa: F401
from .azimuthalaverage import AzimuthalAverage, azimuthal_average  # noqa: F401
from .utils import ddm_files, images2numpy, read_images, tiff2numpy  # noqa: F401
from . import intermediatescatteringfunction, lags, noise_est, sfio, utils, window  # noqa: F401

__version__ = "0.1.0"
