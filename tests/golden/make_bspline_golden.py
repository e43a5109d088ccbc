"""Writes bspline_golden.npz: collocation operators from scipy.interpolate
(an implementation independent of the oracle's de Boor recurrences) for the
basis the reference builds in LABS_fill_full (labs_fill_full.cpp:35-195):
order-k B-splines on breakpoints mapped as the reference does, k-fold end
knots, collocation at Greville abscissae, integration weights = integrals of
the basis functions, centre-line values at y = 0.

Run: python tests/golden/make_bspline_golden.py
"""
import os

import numpy as np
from scipy.interpolate import BSpline

CASES = [
    # (name, ny, gsl order k, map, stretch)   map: 0 uniform/sin rule, 1 tanh
    ("u32k7", 32, 7, 0, 0.0),     # test_bspline.cpp:6-8, uniform
    ("s32k7", 32, 7, 0, 0.99),    # test_bspline.cpp:6-8, sin map 0.99
    ("t33k8", 33, 8, 1, 2.0),     # tanh map, stretch 2.0, order 8
    ("t65k8", 65, 8, 1, 2.0),     # the Re_tau=180 wall-normal grid
]


def breakpoints(nbreak, kmap, stretch, ly=2.0):
    i = np.arange(nbreak, dtype=float)
    xi = -1.0 + (2.0 * i) / (nbreak - 1)
    if kmap == 1:
        return 0.5 * ly * np.tanh(stretch * xi) / np.tanh(stretch)
    if abs(stretch) <= 1e-16:
        return -1.0 + 2.0 / (nbreak - 1) * i
    return np.sin(stretch * np.pi / 2.0 * xi) / np.sin(stretch * np.pi / 2.0)


def case(ny, k, kmap, stretch):
    nb = ny - k + 2
    brk = breakpoints(nb, kmap, stretch)
    t = np.concatenate([np.full(k - 1, brk[0]), brk, np.full(k - 1, brk[-1])])
    y = np.array([t[i + 1:i + k].mean() for i in range(ny)])
    D = np.zeros((5, ny, ny))
    integ = np.zeros(ny)
    center = np.zeros(ny)
    for c in range(ny):
        coef = np.zeros(ny)
        coef[c] = 1.0
        b = BSpline(t, coef, k - 1, extrapolate=True)
        for m in range(5):
            D[m, :, c] = b(y, nu=m) if m <= k - 1 else 0.0
        integ[c] = b.integrate(-1.0, 1.0)
        center[c] = b(0.0)
    # the right end point belongs to the last interval (value 1 for the last basis)
    return dict(brk=brk, y=y, D=D, integ=integ, center=center)


if __name__ == "__main__":
    out = {}
    for name, ny, k, kmap, st in CASES:
        r = case(ny, k, kmap, st)
        for key, val in r.items():
            out[f"{name}_{key}"] = val
        out[f"{name}_params"] = np.array([ny, k, kmap, st])
    x, w = np.polynomial.legendre.leggauss(4)
    out["gl4_x"], out["gl4_w"] = x, w
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "bspline_golden.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")
