"""The Hilbert cell order (hydro_hilbert_order) against the curve's defining
property: on an 8 x 8 grid of centroids, consecutive cells of the order are
grid neighbours (the curve moves one cell at a time), it starts in a corner,
and the order is a permutation; non-finite coordinates are rejected."""
import ctypes

import numpy as np


def test_hilbert_order_walks_the_grid():
    import hydro_amd._lib as L

    lib = L.load()
    n = 8
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    rng = np.random.default_rng(3)
    shuffle = rng.permutation(n * n)
    xc = np.ascontiguousarray((i.ravel() * 10.0 + 3.0)[shuffle])
    yc = np.ascontiguousarray((j.ravel() * 10.0 - 7.0)[shuffle])
    perm = np.zeros(n * n, np.int32)
    p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    assert lib.hydro_hilbert_order(n * n, p(xc), p(yc), p(perm)) == L.OK
    assert np.array_equal(np.sort(perm), np.arange(n * n))
    gx, gy = (xc[perm] - 3.0) / 10.0, (yc[perm] + 7.0) / 10.0
    steps = np.abs(np.diff(gx)) + np.abs(np.diff(gy))
    assert np.all(steps == 1.0)
    assert (gx[0], gy[0]) in {(0, 0), (0, n - 1), (n - 1, 0), (n - 1, n - 1)}
    xc[5] = np.nan
    assert lib.hydro_hilbert_order(n * n, p(xc), p(yc), p(perm)) == L.ERR_INVALID
