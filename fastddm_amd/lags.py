"""Integer lag generators.

``logspace_int`` follows the reference (src/python/lags.py:38-103): a
geometric progression starting at 1 whose ratio is re-derived whenever two
consecutive samples would round to the same integer (the sample is then forced
one step up), so every returned value is distinct. ``fibonacci`` follows
src/python/lags.py:106-150 (always starts with 1, 2).
"""

import numpy as np


def logspace_int(stop: int, num: int = 50, endpoint: bool = False) -> np.ndarray:
    """Quasi-evenly log-spaced distinct integers in ``[1, stop]``.

    With ``endpoint=False`` the progression is laid out for ``num + 1``
    samples and only the first ``num`` are returned, so ``stop`` is excluded.

    >>> logspace_int(10, num=5, endpoint=True)
    array([ 1,  2,  3,  6, 10])
    >>> logspace_int(10, num=5)
    array([1, 2, 3, 4, 7])
    """
    steps = num if endpoint else num + 1
    ratio = float(stop) ** (1.0 / steps)
    vals = np.empty(num, dtype=np.float64)
    vals[0] = 1.0
    for i in range(1, num):
        nxt = vals[i - 1] * ratio
        if nxt - vals[i - 1] >= 1.0:
            vals[i] = nxt
        else:
            vals[i] = vals[i - 1] + 1.0
            ratio = (stop / vals[i]) ** (1.0 / (steps - i - 1))
    return np.round(vals).astype(int)


def fibonacci(stop: int, endpoint: bool = False) -> np.ndarray:
    """Fibonacci samples 1, 2, 3, 5, ... up to ``stop`` (excluded unless ``endpoint``)."""
    limit = stop if endpoint else stop - 1
    out = [1, 2]
    while out[-1] + out[-2] <= limit:
        out.append(out[-1] + out[-2])
    return np.array(out, dtype=int)
