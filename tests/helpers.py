"""Small helpers shared by the tests."""
import numpy as np


def band_geom(n, bw):
    first = lambda r: 0 if r < bw else (n - 1 - 2 * bw if r >= n - bw else r - bw)
    last = lambda r: 2 * bw if r < bw else (n - 1 if r >= n - bw else r + bw)

    def slot(r, c):
        if r < bw and c > r + bw:
            return 2 * bw - c
        if r >= n - bw and c < r - bw:
            return n - 1 - c
        return bw + c - r
    return first, last, slot


def unpack_band(band, bw):
    """Dense matrix from the reference's near-boundary packed band
    (module_labs.f90:301-328)."""
    n = band.shape[0]
    first, last, slot = band_geom(n, bw)
    F = np.zeros((n, n))
    for r in range(n):
        for c in range(first(r), last(r) + 1):
            F[r, c] = band[r, slot(r, c)]
    return F


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)
