"""TEST INFRASTRUCTURE ONLY -- FP64 NumPy oracle of the image structure function.

Only tests/, the smoke check and bench.py's cpu_baseline leg import this
module, as the checker; the fastddm_amd product path never does.

Follows the reference's NumPy core (src/python/_ddm_python.py):

* ``spectra``            <- ``normalized_rfft2``  _ddm_python.py:248-289
* ``autocorrelation``    <- ``autocorrelation``   _ddm_python.py:17-48 (time FFT
  zero-padded to ``nt``; any nt >= 2L-1 gives the same values)
* ``ddm_fft``            <- ``image_structure_function`` + cumsums,
  _ddm_python.py:101-153, 224-238
* ``ddm_diff``           <- ``_diff_image_structure_function``  :51-98
* PS / VAR               <- ``square_mod.mean`` / ``rfft2.var``  :241-242
* output shift           <- ``scipy.fft.fftshift(axes=-2)``     :244

``ddm_fft(..., subtract_mean=True)`` removes the temporal mean of every q
before the correlation; D is unchanged in exact arithmetic and the FP64
cancellation at low |q| is reduced, which makes it the reference used for
precision reporting. ``brute_force_D`` evaluates the defining sum directly.

Pinning (tests/test_oracle.py): numpy.fft against brute-force DFT sums,
fft == diff == brute force, nt independence, constant / impulse / plane-wave
closed forms, and agreement with the independent C restatement
(oracle/ddm_port.c).
"""

import numpy as np


def spectra(img, nx, ny, window=None):
    """Normalised half-plane spectra (L, ny, nx//2+1), complex128."""
    x = np.asarray(img).astype(np.float64)
    if window is not None and np.size(window) > 0:
        x = x * np.asarray(window, dtype=np.float64)
    return np.fft.rfft2(x, s=(ny, nx)) / np.sqrt(nx * ny)


def autocorrelation(F, nt):
    """Re C(dt) = Re sum_t conj(F(t)) F(t+dt), dt = 0..L-1, along axis 0."""
    L = F.shape[0]
    G = np.fft.fft(F, n=nt, axis=0)
    return np.fft.ifft(np.abs(G) ** 2, axis=0)[:L].real


def _shift(out):
    return np.fft.fftshift(out, axes=-2)


def ddm_fft(img, lags, nx, ny, nt, window=None, subtract_mean=False):
    """fft-mode result (NL+2, ny, nx//2+1) float64, y-shifted."""
    return _shift(structure_from_spectra(spectra(img, nx, ny, window), lags, nt, subtract_mean))


def structure_from_spectra(F, lags, nt, subtract_mean=False):
    """D slices, PS and VAR (NL+2, *F.shape[1:]) of spectra F (L, ...), unshifted.

    The time stage of ``ddm_fft`` on any subset of q (columns of F)."""
    lags = np.asarray(lags, dtype=np.int64)
    L = F.shape[0]
    ps = np.mean(np.abs(F) ** 2, axis=0)
    var = np.var(F, axis=0)
    G = F - F.mean(axis=0) if subtract_mean else F
    sq = np.abs(G) ** 2
    cs = np.cumsum(sq, axis=0)
    cs_rev = np.cumsum(sq[::-1], axis=0)
    ac = autocorrelation(G, nt)
    out = np.empty((len(lags) + 2,) + F.shape[1:], dtype=np.float64)
    for i, lag in enumerate(lags):
        off = lag + 1
        out[i] = (cs[-off] + cs_rev[-off] - 2.0 * ac[lag]) / (L - lag)
    out[-2] = ps
    out[-1] = var
    return out


def ddm_diff(img, lags, nx, ny, window=None):
    """diff-mode result (NL+2, ny, nx//2+1) float64, y-shifted."""
    lags = np.asarray(lags, dtype=np.int64)
    F = spectra(img, nx, ny, window)
    out = np.empty((len(lags) + 2,) + F.shape[1:], dtype=np.float64)
    for i, lag in enumerate(lags):
        d = F[lag:] - F[: F.shape[0] - lag]
        out[i] = np.mean(np.abs(d) ** 2, axis=0)
    out[-2] = np.mean(np.abs(F) ** 2, axis=0)
    out[-1] = np.var(F, axis=0)
    return _shift(out)


def brute_force_D(F, lag):
    """Direct (1/(L-lag)) sum_t |F(t+lag) - F(t)|^2 with explicit loops over t."""
    L = F.shape[0]
    acc = np.zeros(F.shape[1:], dtype=np.float64)
    for t in range(L - lag):
        acc += np.abs(F[t + lag] - F[t]) ** 2
    return acc / (L - lag)


def brute_force_dft2(frame, nx, ny):
    """Direct O(N^2) 2D DFT of one zero-padded real frame, half plane, normalised."""
    H, W = frame.shape
    x = np.zeros((ny, nx))
    x[:H, :W] = frame
    nxh = nx // 2 + 1
    ey = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    ex = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nxh)) / nx)
    return (ey @ x @ ex) / np.sqrt(nx * ny)


def relative_error(x, ref):
    """|x - ref| / |ref| over cells where ref != 0 (flattened)."""
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    m = ref != 0
    return np.abs(x[m] - ref[m]) / np.abs(ref[m])


def azimuthal_average_array(data, dist, bin_edges, w, w2, eval_err=True):
    """Weighted azimuthal average over pre-digitised pixels (FP64).

    Restates azimuthal_average_array of the reference
    (src/python/azimuthalaverage.py:420-480) for one flattened pixel set:
    ``w``/``w2`` are counts*weights and counts*weights**2 with zeros for the
    excluded pixels, bins come from ``np.searchsorted(bin_edges, dist)``.
    Returns (avg, err) of shape (nbins, rows): avg = sum w x / sum w, NaN for
    empty bins; err = sqrt(sum w (x - avg)^2 / (sum w - sum w2 / sum w)).
    """
    nb = len(bin_edges)
    rows = data.shape[0]
    x = data.reshape(rows, -1).astype(np.float64)
    idx = np.minimum(np.searchsorted(bin_edges, dist.ravel()), nb - 1)
    wf, w2f = w.ravel().astype(np.float64), w2.ravel().astype(np.float64)
    sw = np.bincount(idx, weights=wf, minlength=nb)
    sw2 = np.bincount(idx, weights=w2f, minlength=nb)
    onehot = np.zeros((nb, idx.size))
    onehot[idx, np.arange(idx.size)] = wf
    with np.errstate(divide="ignore", invalid="ignore"):
        avg = (onehot @ x.T) / sw[:, None]
        err = None
        if eval_err:
            dev = (x.T[None, :, :] - avg[:, None, :]) ** 2          # (nb, npix, rows)
            num = np.einsum("bp,bpr->br", onehot, dev)
            bias = np.where(sw > 0, sw - sw2 / np.where(sw > 0, sw, 1.0), 0.0)
            err = np.sqrt(num / bias[:, None])
            err[sw == 0] = np.nan
    avg[sw == 0] = np.nan
    return avg, err
