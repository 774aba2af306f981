"""Padded FFT lengths per core.

Behaviour follows ``next_fast_len`` in the reference (src/python/_fftopt.py:29-150):

* a target that is already a power of two is returned as is (``:76-77``);
* ``'py'`` defers to ``scipy.fft.next_fast_len`` (``:78-79``);
* ``'cpp'`` returns the smallest 2^a 3^b 5^c 7^d, optionally times one factor
  of 11 or 13, that is >= target (``:80-115``);
* ``'cuda'`` returns the smallest 2^a 3^b 5^c 7^d >= target, optionally even
  (``:116-148``);
* ``'hip'`` (new) uses the ``'cuda'`` rule: the HIP FFT kernels implement
  radices 2, 3, 5 and 7;
* any other core raises ``ValueError`` (``:150``).

The search here is written differently (a direct walk over candidates), but
returns the same integers; tests/test_helpers.py checks it against a brute-force
enumeration of smooth numbers for every core.
"""

from typing import Optional

from scipy import fft as _scipy_fft


def next_power_of_2(target: int) -> int:
    """Smallest power of two >= ``target`` (reference ``_fftopt.py:12-26``)."""
    return 1 << (int(target) - 1).bit_length()


def _strip(n: int, primes) -> int:
    for p in primes:
        while n % p == 0:
            n //= p
    return n


def _is_smooth(n: int, extra: bool) -> bool:
    """True if n = 2^a 3^b 5^c 7^d (times at most one 11 or 13 when ``extra``)."""
    r = _strip(n, (2, 3, 5, 7))
    if r == 1:
        return True
    return extra and r in (11, 13)


def next_fast_len(
    target: int, core: Optional[str] = "py", force_even: Optional[bool] = False
) -> int:
    """Next fast FFT length >= ``target`` for the given backend core.

    Parameters
    ----------
    target : int
        Target length (>= 1).
    core : str, optional
        'py', 'cpp', 'cuda' or 'hip'. Default is 'py'.
    force_even : bool, optional
        Only with 'cuda'/'hip': return an even length.

    Raises
    ------
    ValueError
        If ``core`` is not supported.
    """
    target = int(target)
    max_val = next_power_of_2(target)
    if target == max_val:
        return max_val
    name = core.upper()
    if name == "PY":
        return _scipy_fft.next_fast_len(target)
    if name in ("CPP", "CUDA", "HIP"):
        extra = name == "CPP"
        even = force_even and name != "CPP"
        n = target
        while n < max_val:
            if _is_smooth(n, extra) and (not even or n % 2 == 0):
                return n
            n += 1
        return max_val
    raise ValueError("Mode not supported in `next_fast_len`.")
