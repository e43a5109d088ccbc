"""Pins of the oracle's statistics-time pressure (get_pressure.f90,
get_statistics.f90:689-716 and 982-1017).

Closed forms: with polynomial data of degree < order the B-spline
collocation operators are exact, so the pressure Poisson solve
(D2 - k^2 D0) c = rhs with Neumann wall rows D1 c = bc returns the
manufactured pressure at the collocation points to roundoff."""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from tests.helpers import unpack_band


def setup(oracle_lib, name="tiny", **kw):
    cfg = dataclasses.replace(preset(name), **kw)
    orc = oracle_lib.Oracle(cfg)
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
    P = np.zeros((10, cfg.nz, cfg.nxh, cfg.ny), np.complex128)
    kx = np.arange(cfg.nxh) / cfg.lmfx
    nh = cfg.nz // 2
    kzi = np.concatenate([np.arange(nh + 1), np.arange(nh + 2, cfg.nz + 1) - 2 * nh - 1])
    return cfg, orc, S, P, kx, kzi / cfg.lmfz


def q4(y):  # (1 - y^2)^2 and its second derivative
    return (1 - y * y) ** 2, 12 * y * y - 4


@pytest.mark.parametrize("term", ["u1", "u3", "u5", "u6"])
def test_manufactured_pressure(oracle_lib, term):
    cfg, orc, S, P, kx, kz = setup(oracle_lib)
    y = orc.y
    j, k = 2, 3                      # one (kx, kz) line, kx and kz != 0
    a, b = kx[k], kz[j]
    k2 = a * a + b * b
    p, d2p = q4(y)
    lap = d2p - k2 * p                # p'' - k^2 p, p'(+-1) = 0
    if term == "u1":
        P[0, j, k] = lap / (a * a)
        want = p
    elif term == "u3":
        P[2, j, k] = lap / (b * b)
        want = p
    elif term == "u5":
        P[4, j, k] = lap / (2 * a * b)
        want = p
    else:                             # k^2 r - r'' with r = p: solution -p
        P[5, j, k] = p
        want = -p
    Pin = P.copy()
    vv = orc.pressure(S, P)
    assert np.abs(P[5, j, k] - want).max() <= 1e-11 * np.abs(want).max()
    others = np.ones(P.shape[1:3], bool)
    others[j, k] = False
    assert not np.any(P[5][others])   # zero data, zero pressure (and <vv> = 0)
    assert not np.any(vv)
    for f in (0, 1, 2, 3, 4, 6, 7, 8, 9):
        assert np.array_equal(P[f], Pin[f])  # the products are only read


def test_wall_values_from_v(oracle_lib):
    """bc = nu v''(wall): B-spline second derivatives at a clamped wall
    involve only the first three coefficients, so the reference's three-term
    rows are exact for polynomial v; with zero products p solves
    p'' - k^2 p = 0, p'(+-1) = 8 nu, i.e. p = 8 nu sinh(k y) / (k cosh k)."""
    cfg, orc, S, P, kx, kz = setup(oracle_lib, "re180_64")
    y = orc.y
    j, k = 1, 1
    k2 = kx[k] ** 2 + kz[j] ** 2
    kk = np.sqrt(k2)
    S[1, j, k] = (1 - y * y) ** 2 * (0.3 - 0.7j)   # v'' = (12 y^2 - 4) c: 8c at both walls
    vv = orc.pressure(S, P)
    want = 8 * cfg.nu * (0.3 - 0.7j) * np.sinh(kk * y) / (kk * np.cosh(kk))
    assert np.abs(P[5, j, k] - want).max() <= 1e-6 * np.abs(want).max()   # discretisation of sinh
    # mean line: p = -<vv>, <vv> = 2 |v|^2 summed over the kx > 0 modes
    assert np.allclose(vv, 2 * np.abs(S[1, j, k]) ** 2, rtol=1e-15, atol=0)
    assert np.array_equal(P[5, 0, 0], -vv + 0j)


def test_dense_restatement(oracle_lib):
    """Random products on every line against a dense LAPACK solve of the same
    equations assembled from the unpacked operators (including the
    reference's reuse of the wall coefficients of the first solve)."""
    cfg, orc, S, P, kx, kz = setup(oracle_lib, "tiny")
    rng = np.random.default_rng(3)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    P[:] = rng.standard_normal(P.shape) + 1j * rng.standard_normal(P.shape)
    Pin = P.copy()
    orc.pressure(S, P)
    bw, n = orc.bw, cfg.ny
    D0, D1, D2 = (unpack_band(m, bw) for m in (orc.d0, orc.d1, orc.d2))
    iD0 = np.linalg.inv(D0)
    worst = 0.0
    for j in range(cfg.nz):
        for k in range(cfg.nxh):
            if k == 0 and j == 0:
                continue
            a, b = kx[k], kz[j]
            k2 = a * a + b * b
            u1, u2, u3, u4, u5, u6 = (Pin[f, j, k] for f in range(6))
            c6 = iD0 @ u6
            w = 1j * a * u2 + 1j * b * u4
            w[0], w[-1] = c6[0], c6[-1]
            rhs = a * a * u1 + b * b * u3 + k2 * u6 + 2 * a * b * u5 - D2 @ c6 - 2 * (D1 @ (iD0 @ w))
            cv = iD0 @ S[1, j, k]
            rhs[0] = cfg.nu * (D2[0, :3] @ cv[:3])
            rhs[-1] = cfg.nu * (D2[-1, -3:] @ cv[-3:])
            A = D2 - k2 * D0
            A[0], A[-1] = D1[0], D1[-1]
            p = D0 @ np.linalg.solve(A, rhs)
            worst = max(worst, np.abs(p - P[5, j, k]).max() / np.abs(p).max())
    assert worst <= 1e-9, worst
