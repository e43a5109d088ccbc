"""Pins of the oracle's statistics profiles (get_statistics.f90:54-155,
982-1048, 546-652): closed forms under collocation and Parseval."""
import dataclasses

import numpy as np

from channel_passive_amd.config import preset


def make(oracle_lib, name="tiny", **kw):
    cfg = dataclasses.replace(preset(name), **kw)
    orc = oracle_lib.Oracle(cfg)
    return cfg, orc, np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), np.complex128)


def test_mean_profiles_and_gradients(oracle_lib):
    """Poiseuille-like polynomial means: U = 1.5(1-y^2), W = 0.2 y, Theta =
    (1-y^2)^2 are exact under collocation, so the gradients are -3y, 0.2 and
    -4y(1-y^2), the centre values 1.5 (+u_gal) and 1."""
    cfg, orc, S = make(oracle_lib, u_gal=0.25)
    y = orc.y
    S[0, 0, 0] = 1.5 * (1 - y * y)
    S[2, 0, 0] = 0.2 * y
    S[3, 0, 0] = (1 - y * y) ** 2
    st = orc.statistics(S)
    assert np.allclose(st["U"], 1.5 * (1 - y * y) + 0.25, rtol=0, atol=1e-15)
    assert np.array_equal(st["W"], 0.2 * y)
    assert np.allclose(st["dUdy"], -3 * y, rtol=0, atol=1e-12)
    assert np.allclose(st["dWdy"], 0.2, rtol=0, atol=1e-12)
    assert np.allclose(st["dTdy"], -4 * y * (1 - y * y), rtol=0, atol=1e-12)
    assert abs(st["u_center"] - 1.75) < 1e-13
    assert abs(st["t_center"] - 1.0) < 1e-13
    # only the mean mode: <uu> = U^2 of the stored mode, <uv> = 0
    assert np.allclose(st["uu"], (1.5 * (1 - y * y)) ** 2, rtol=1e-15, atol=0)
    assert not np.any(st["uv"])


def test_parseval_single_modes(oracle_lib):
    """u = a(y) cos(kx x) stored as u_k = a/2 at kx > 0: <uu> = a^2/2 (weight
    2 for kx > 0); a kz-only mode on the kx = 0 plane with its conjugate at
    -kz counts once per stored mode; <ut> of u = a cos, theta = b sin is 0."""
    cfg, orc, S = make(oracle_lib)
    y = orc.y
    a, b = 1 - y * y, y * (1 - y * y)
    S[0, 0, 1] = a / 2            # u = a cos(x)
    S[3, 0, 1] = -1j * b / 2      # theta = b sin(x)
    S[2, 1, 0] = 0.3 * a          # w mode (kx = 0, kz = +1) ...
    S[2, cfg.nz - 1, 0] = 0.3 * a  # ... and its conjugate at kz = -1
    st = orc.statistics(S)
    assert np.allclose(st["uu"], a * a / 2, rtol=1e-15, atol=1e-18)
    assert np.allclose(st["tt"], b * b / 2, rtol=1e-15, atol=1e-18)
    assert np.allclose(st["ww"], 2 * (0.3 * a) ** 2, rtol=1e-15, atol=1e-18)
    assert np.allclose(st["ut"], 0.0, atol=1e-18)
    assert not np.any(st["vv"]) and not np.any(st["uv"])


def test_product_order_and_weights(oracle_lib):
    """Random state: every profile equals the Parseval sum written with numpy
    (kx = 0 once, kx > 0 twice) to roundoff, and uv is symmetric in its
    arguments (vt = tv, ...)."""
    cfg, orc, S = make(oracle_lib)
    rng = np.random.default_rng(4)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    st = orc.statistics(S)
    w = np.full(cfg.nxh, 2.0)
    w[0] = 1.0
    pairs = {"uu": (0, 0), "vv": (1, 1), "ww": (2, 2), "tt": (3, 3), "uv": (0, 1), "uw": (0, 2), "vw": (1, 2),
             "ut": (0, 3), "wt": (3, 2), "vt": (1, 3)}
    for name, (p, q) in pairs.items():
        want = np.einsum("k,jky->y", w, (S[p].conj() * S[q]).real)
        assert np.allclose(st[name], want, rtol=1e-12, atol=0), name
