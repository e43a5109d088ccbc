"""Pins of the oracle's B-spline operators (labs_fill_full.cpp:35-256,
module_labs.f90:245-328) and of the product's own generator.

* the reference's only numerical test, tests/test_bspline.cpp:5-27
  (partition of unity, derivative row sums, weights sum to Ly = 2)
* scipy.interpolate golden fixtures (tests/golden/bspline_golden.npz)
* closed forms: exact differentiation and integration of polynomials of
  degree < order
* bandwidth LABS_BW and the packed-band pattern
"""
import ctypes
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "bspline_golden.npz"))


@pytest.mark.parametrize("stretch", [0.0, 0.99])
def test_reference_bspline_test(oracle_lib, stretch):
    """test_bspline.cpp: k=7, n=32, Ly=2, stretch in {0, 0.99}."""
    r = oracle_lib.labs_setup(32, 6, 0, stretch, 2.0)
    assert np.isfinite(r["integ"]).all()
    assert abs(r["integ"].sum() - 2.0) <= 1e-12
    for d in range(5):
        M = r["full"][d]
        assert np.isfinite(M).all()
        rowsum = M.sum(axis=1)
        scale = 1.0 + np.abs(M).sum(axis=1)
        target = 1.0 if d == 0 else 0.0
        assert np.all(np.abs(rowsum - target) <= 1e-11 * scale), d


@pytest.mark.parametrize("name", ["u32k7", "s32k7", "t33k8", "t65k8"])
def test_against_scipy(oracle_lib, name):
    ny, k, kmap, st = GOLD[f"{name}_params"]
    ny, k, kmap = int(ny), int(k), int(kmap)
    r = oracle_lib.labs_setup(ny, k - 1, kmap, float(st), 2.0)
    assert np.allclose(r["brk"], GOLD[f"{name}_brk"], rtol=0, atol=1e-15)
    assert np.allclose(r["y"], GOLD[f"{name}_y"], rtol=0, atol=4.5e-16)  # mean vs running mean: <= 2 ulp
    for d in range(min(5, k)):
        M, ref = r["full"][d], GOLD[f"{name}_D"][d]
        err = np.abs(M - ref).max() / np.abs(ref).max()
        assert err <= 1e-13, (d, err)
    assert np.allclose(r["integ"], GOLD[f"{name}_integ"], rtol=1e-13, atol=1e-16)
    assert np.allclose(r["center"], GOLD[f"{name}_center"], rtol=0, atol=1e-15)


@pytest.mark.parametrize("ny,labs_k,kmap,stretch", [(65, 7, 1, 2.0), (129, 7, 1, 2.0), (33, 6, 0, 0.9)])
def test_polynomials_exact(oracle_lib, ny, labs_k, kmap, stretch):
    """Collocation at Greville points reproduces polynomials of degree < order:
    D1 D0^-1 y^p = p y^(p-1), D2 D0^-1 y^p = p(p-1) y^(p-2), and the weights
    integrate them exactly."""
    r = oracle_lib.labs_setup(ny, labs_k, kmap, stretch, 2.0)
    y = r["y"]
    D0, D1, D2 = r["full"][:3]
    for p in range(labs_k + 1):
        c = np.linalg.solve(D0, y ** p)
        d1 = p * y ** (p - 1) if p >= 1 else 0 * y
        d2 = p * (p - 1) * y ** (p - 2) if p >= 2 else 0 * y
        assert np.allclose(D1 @ c, d1, rtol=0, atol=1e-9 * max(1, p * p))
        assert np.allclose(D2 @ c, d2, rtol=0, atol=1e-7 * max(1, p ** 3))
        exact = (1 - (-1) ** (p + 1)) / (p + 1)
        assert abs(r["integ"] @ c - exact) <= 1e-13


def test_gauss_legendre_rule(oracle_lib):
    """The 4-point rule the integration weights use integrates degree-7
    polynomials on every interval exactly; checked on the whole interval
    against numpy's leggauss."""
    x, w = GOLD["gl4_x"], GOLD["gl4_w"]
    for p in range(8):
        exact = (1 - (-1) ** (p + 1)) / (p + 1)
        assert abs(np.sum(w * x ** p) - exact) <= 1e-15


@pytest.mark.parametrize("ny,labs_k,kmap,stretch,bw", [(65, 7, 1, 2.0, 4), (385, 7, 1, 2.0, 4), (129, 7, 1, 2.0, 4),
                                                       (64, 6, 0, 0.0, 3), (65, 6, 0, 0.0, 4), (32, 6, 0, 0.99, 4)])
def test_bandwidth(oracle_lib, ny, labs_k, kmap, stretch, bw):
    r = oracle_lib.labs_setup(ny, labs_k, kmap, stretch, 2.0)
    assert r["bw"] == bw
    # every entry outside the near-boundary packed pattern is exactly zero
    for M in r["full"][:3]:
        band = oracle_lib.full_to_band(M, bw)
        from tests.helpers import unpack_band
        assert np.array_equal(unpack_band(band, bw), M)
    # and bw is the smallest such width (LABS_eval_bw searches upward from 1)
    lost = [not np.array_equal(__import__("tests.helpers", fromlist=["x"]).unpack_band(
        oracle_lib.full_to_band(M, bw - 1), bw - 1), M) for M in r["full"][:3]]
    assert any(lost)


@pytest.mark.parametrize("name", ["re180_64", "re1000", "tiny"])
def test_product_setup_bitwise(oracle_lib, emu_lib, name):
    """The product's host generator (channel_passive_amd/csrc/bspline.cpp)
    equals the oracle bit for bit, so the device solves see the same matrices."""
    from channel_passive_amd.config import preset
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    n, W = cfg.ny, 2 * orc.bw + 1
    dp = ctypes.POINTER(ctypes.c_double)
    bw = ctypes.c_int(0)
    arrs = [np.zeros(n)] + [np.zeros((n, W)) for _ in range(4)] + [np.zeros(n)]
    c = cfg.to_c()
    assert emu_lib.emu_setup(ctypes.byref(c), ctypes.byref(bw), *[a.ctypes.data_as(dp) for a in arrs]) == 0
    assert bw.value == orc.bw
    for a, b in zip(arrs, (orc.y, orc.d0, orc.d1, orc.d2, orc.d0lu, orc.integ)):
        assert np.array_equal(a, b)
