"""CPU checks of the azimuthal-average oracle and of the host-side bin maps
(fastddm_amd.azimuthalaverage.bin_maps) against per-pixel loops and closed
forms."""

import numpy as np
import pytest

from fastddm_amd.azimuthalaverage import bin_maps
from oracle import oracle_np as onp


def _loops(data, dist, bin_edges, w, w2):
    """Per-pixel loops of the defining sums (the reference's own loop form)."""
    nb, rows = len(bin_edges), data.shape[0]
    x = data.reshape(rows, -1)
    idx = np.minimum(np.searchsorted(bin_edges, dist.ravel()), nb - 1)
    sw, sw2 = np.zeros(nb), np.zeros(nb)
    for p, b in enumerate(idx):
        sw[b] += w.ravel()[p]
        sw2[b] += w2.ravel()[p]
    avg = np.zeros((nb, rows))
    for p, b in enumerate(idx):
        if w.ravel()[p] > 0:
            avg[b] += x[:, p] * w.ravel()[p] / sw[b]
    avg[sw == 0] = np.nan
    err = np.zeros((nb, rows))
    for p, b in enumerate(idx):
        if w.ravel()[p] > 0:
            with np.errstate(divide="ignore", invalid="ignore"):
                err[b] += w.ravel()[p] * (x[:, p] - avg[b]) ** 2 / (sw[b] - sw2[b] / sw[b])
    err[sw == 0] = np.nan
    return avg, np.sqrt(err)


def _grid(ny=12, nx=10):
    kx = 2 * np.pi * np.fft.fftfreq(nx)[: nx // 2 + 1]
    ky = 2 * np.pi * np.fft.fftshift(np.fft.fftfreq(ny))
    X, Y = np.meshgrid(kx, ky)
    return np.sqrt(X ** 2 + Y ** 2), kx, ky


@pytest.mark.parametrize("bins", [5, 9, [0.2, 0.9, 1.5, 2.5, 3.0]])
@pytest.mark.parametrize("use_mask", [False, True])
def test_oracle_matches_pixel_loops(bins, use_mask):
    dist, _, _ = _grid()
    rng = np.random.default_rng(0)
    data = rng.random((4,) + dist.shape) * 10
    counts = np.full(dist.shape, 2)
    counts[:, 0] = 1
    mask = rng.random(dist.shape) > 0.2 if use_mask else None
    weights = rng.random(dist.shape) + 0.5
    bi, w, w2, edges, k = bin_maps(dist, bins, None, mask, weights, counts)
    avg, err = onp.azimuthal_average_array(data, dist, edges, w, w2)
    la, le = _loops(data, dist, edges, w, w2)
    assert np.allclose(avg, la, rtol=1e-12, equal_nan=True)
    assert np.allclose(err, le, rtol=1e-10, equal_nan=True)


def test_constant_data_has_zero_error_and_k_map_averages_to_k():
    dist, _, _ = _grid(16, 14)
    counts = np.full(dist.shape, 2)
    bi, w, w2, edges, k = bin_maps(dist, 6, None, None, None, counts)
    avg, err = onp.azimuthal_average_array(np.full((3,) + dist.shape, 7.5), dist, edges, w, w2)
    full = ~np.isnan(avg[:, 0])
    assert np.allclose(avg[full], 7.5) and np.allclose(err[full], 0.0)
    # averaging the distance map itself gives the bins' k
    avg, _ = onp.azimuthal_average_array(dist[None], dist, edges, w, w2)
    assert np.allclose(avg[full, 0], k[full])


def test_two_pixel_bin_closed_form():
    dist = np.array([[1.0, 2.0]])
    data = np.array([[[3.0, 5.0]]])
    w = np.array([[1.0, 3.0]])
    avg, err = onp.azimuthal_average_array(data, dist, np.array([0.0, 2.0]), w, w ** 2)
    m = (3.0 * 1 + 5.0 * 3) / 4
    var = (1 * (3 - m) ** 2 + 3 * (5 - m) ** 2) / (4 - 10 / 4)
    assert np.isclose(avg[1, 0], m) and np.isclose(err[1, 0], np.sqrt(var))
    assert np.isnan(avg[0, 0]) and np.isnan(err[0, 0])


def test_bin_maps_edge_rules():
    dist = np.array([[0.0, 0.5, 1.0, 1.5, 2.0, 2.5]])
    bi, w, w2, edges, k = bin_maps(dist, 3, (0.0, 2.0), None, None, None)
    # edges 0, 1, 2: searchsorted -> 0, 1, 1, 2, 2, 3 (3 = beyond the last edge: excluded)
    assert list(edges) == [0.0, 1.0, 2.0]
    assert list(bi[0]) == [0, 1, 1, 2, 2, 2]
    assert list(w[0]) == [1, 1, 1, 1, 1, 0]
    assert np.allclose(k, [0.0, 0.75, 1.75])
    with pytest.raises(ValueError):
        bin_maps(dist, [0.0, 2.0, 1.0])
    with pytest.raises(ValueError):
        bin_maps(dist, 3, weights=-np.ones_like(dist))
    with pytest.warns(UserWarning):
        bin_maps(dist, 3, mask=np.ones_like(dist))


def test_device_maps_cache():
    """Bin maps are kept per geometry: the same inputs return the same
    (read-only) arrays, a different bin count or mask builds new ones, and a
    non-boolean mask is never cached (its cast warns on every call)."""
    from fastddm_amd import azimuthalaverage as az

    az._MAPS_CACHE.clear()
    _, kx, ky = _grid(16, 14)
    a = az._device_maps(kx, ky, 14, 16, 8, 10, None, None, None, True)
    b = az._device_maps(kx.copy(), ky.copy(), 14, 16, 8, 10, None, None, None, True)
    assert all(x is y for x, y in zip(a, b))
    assert not a[0].flags.writeable and not a[1].flags.writeable
    c = az._device_maps(kx, ky, 14, 16, 8, 12, None, None, None, True)
    assert len(c[3]) == 12 and c[0] is not a[0]
    mask = np.ones((16, 8), dtype=bool)
    mask[3, 2] = False
    d = az._device_maps(kx, ky, 14, 16, 8, 10, None, mask, None, True)
    assert d[1][3, 2] == 0 and a[1][3, 2] != 0
    with pytest.warns(UserWarning):
        az._device_maps(kx, ky, 14, 16, 8, 10, None, mask.astype(int), None, True)
    with pytest.warns(UserWarning):
        az._device_maps(kx, ky, 14, 16, 8, 10, None, mask.astype(int), None, True)
    for i in range(2 * az._MAPS_CACHE_SIZE):
        az._device_maps(kx, ky, 14, 16, 8, 20 + i, None, None, None, True)
    assert len(az._MAPS_CACHE) == az._MAPS_CACHE_SIZE
    # the maps equal a fresh build
    dist = np.sqrt(np.meshgrid(kx, ky)[0] ** 2 + np.meshgrid(kx, ky)[1] ** 2)
    fresh = bin_maps(dist, 10, None, None, None, az._half_plane_counts((16, 8), 14), True)
    for x, y in zip(a, fresh):
        assert np.array_equal(x, y)
