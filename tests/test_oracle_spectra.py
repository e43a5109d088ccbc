"""Pins of the oracle's 1-D energy (co)spectra (write_energy_spectra.f90:9-300,
326-399; get_y_dir_deriv_of_field and get_vorticity_field,
get_statistics.f90:658-681, 912-948): single Fourier modes with polynomial
y-profiles (derivatives exact under collocation), Parseval against the
profile sums, and a dense numpy restatement on a random state."""
import dataclasses

import numpy as np

from channel_passive_amd.config import physical_wavenumbers, preset
from tests.helpers import unpack_band

NAMES = ("uu", "vv", "ww", "tt", "uv", "vw", "uw", "ut", "vt", "wt", "oxox", "oyoy", "ozoz", "oxoy", "oxoz",
         "oyoz", "voz", "woy")


def make(oracle_lib, name="tiny", **kw):
    cfg = dataclasses.replace(preset(name), **kw)
    orc = oracle_lib.Oracle(cfg)
    return cfg, orc, np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), np.complex128)


def test_names_match_oracle(oracle_lib):
    assert oracle_lib.Oracle.SPECTRA_NAMES == NAMES


def test_single_mode_polynomial_profiles(oracle_lib):
    """One mode (kx index 2, kz index 1) with u = 1 - y^2, v = 0.5 (1 - y^2)^2,
    w = y (1 - y^2): du/dy = -2y and dw/dy = 1 - 3y^2 exactly, so
    ox = (1 - 3y^2) - i kz v, oy = i (kz u - kx w), oz = i kx v + 2y; every
    spectrum lives in that single kx (kz) bin with weight 2 (kx > 0)."""
    cfg, orc, S = make(oracle_lib)
    y = orc.y
    kxi, kzi, kxw, kzw = physical_wavenumbers(cfg)
    j, k = 1, 2
    kx, kz = kxw[k], kzw[j]
    u, v, w = 1 - y * y, 0.5 * (1 - y * y) ** 2, y * (1 - y * y)
    S[0, j, k], S[1, j, k], S[2, j, k] = u, v, w
    ex, ez = orc.spectra(S)
    ox = (1 - 3 * y * y) - 1j * kz * v
    oy = 1j * (kz * u - kx * w)
    oz = 1j * kx * v + 2 * y
    fields = {"u": u + 0j, "v": v + 0j, "w": w + 0j, "t": 0 * y + 0j, "ox": ox, "oy": oy, "oz": oz}
    split = {"uu": ("u", "u"), "vv": ("v", "v"), "ww": ("w", "w"), "tt": ("t", "t"), "uv": ("u", "v"),
             "vw": ("v", "w"), "uw": ("u", "w"), "ut": ("u", "t"), "vt": ("v", "t"), "wt": ("w", "t"),
             "oxox": ("ox", "ox"), "oyoy": ("oy", "oy"), "ozoz": ("oz", "oz"), "oxoy": ("ox", "oy"),
             "oxoz": ("ox", "oz"), "oyoz": ("oy", "oz"), "voz": ("v", "oz"), "woy": ("w", "oy")}
    for q, name in enumerate(NAMES):
        a, b = split[name]
        want = 2 * (fields[a].conj() * fields[b]).real
        assert np.allclose(ex[q, k], want, rtol=0, atol=1e-11), name
        mask = np.ones(cfg.nxh, bool)
        mask[k] = False
        assert not np.any(ex[q][mask]), name
        # the kz list: the reference's "ozoz" dataset is (ox, oz)
        za, zb = ("ox", "oz") if name == "ozoz" else (a, b)
        assert np.allclose(ez[q, j], 2 * (fields[za].conj() * fields[zb]).real, rtol=0, atol=1e-11), name
        zmask = np.ones(cfg.nz, bool)
        zmask[j] = False
        assert not np.any(ez[q][zmask]), name


def test_quirk_and_parseval(oracle_lib):
    """Random state: the kz "ozoz" dataset is the (ox, oz) co-spectrum
    (identical to "oxoz"); every velocity/scalar spectrum sums over its
    wavenumber to the Parseval profile of get_statistics (same terms, other
    order); the kx = 0 bin carries weight 1."""
    cfg, orc, S = make(oracle_lib)
    rng = np.random.default_rng(8)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    ex, ez = orc.spectra(S)
    iz = {n: i for i, n in enumerate(NAMES)}
    assert np.array_equal(ez[iz["ozoz"]], ez[iz["oxoz"]])
    st = orc.statistics(S)
    for name, stat in (("uu", "uu"), ("vv", "vv"), ("ww", "ww"), ("tt", "tt"), ("uv", "uv"), ("vw", "vw"),
                       ("uw", "uw"), ("ut", "ut"), ("vt", "vt"), ("wt", "wt")):
        q = iz[name]
        assert np.allclose(ex[q].sum(axis=0), st[stat], rtol=1e-12, atol=0), name
        assert np.allclose(ez[q].sum(axis=0), st[stat], rtol=1e-12, atol=0), name
    want0 = (S[0, :, 0].conj() * S[0, :, 0]).real.sum(axis=0)
    assert np.allclose(ex[iz["uu"], 0], want0, rtol=1e-13, atol=0)


def test_dense_restatement(oracle_lib):
    """Random state against numpy: du/dy = D1 D0^-1 u with dense matrices
    from the packed bands, vorticity, weighted sums (roundoff only)."""
    cfg, orc, S = make(oracle_lib)
    rng = np.random.default_rng(9)
    S[:] = rng.standard_normal(S.shape) + 1j * rng.standard_normal(S.shape)
    ex, ez = orc.spectra(S)
    D0, D1 = unpack_band(orc.d0, orc.bw), unpack_band(orc.d1, orc.bw)
    Dy = D1 @ np.linalg.inv(D0)
    _, _, kxw, kzw = physical_wavenumbers(cfg)
    u, v, w, t = S
    du, dw = u @ Dy.T, w @ Dy.T
    KX, KZ = kxw[None, :, None], kzw[:, None, None]
    ox = dw - 1j * KZ * v
    oy = 1j * KZ * u - 1j * KX * w
    oz = 1j * KX * v - du
    F = {"u": u, "v": v, "w": w, "t": t, "ox": ox, "oy": oy, "oz": oz}
    pairs = [("u", "u"), ("v", "v"), ("w", "w"), ("t", "t"), ("u", "v"), ("v", "w"), ("u", "w"), ("u", "t"),
             ("v", "t"), ("w", "t"), ("ox", "ox"), ("oy", "oy"), ("oz", "oz"), ("ox", "oy"), ("ox", "oz"),
             ("oy", "oz"), ("v", "oz"), ("w", "oy")]
    wx = np.full(cfg.nxh, 2.0)
    wx[0] = 1.0
    for q, (a, b) in enumerate(pairs):
        prod = (F[a].conj() * F[b]).real * wx[None, :, None]
        scale = np.abs(prod).sum(axis=0).max()
        assert np.allclose(ex[q], prod.sum(axis=0), rtol=0, atol=1e-12 * scale), NAMES[q]
        za, zb = ("ox", "oz") if NAMES[q] == "ozoz" else (a, b)
        prodz = (F[za].conj() * F[zb]).real * wx[None, :, None]
        assert np.allclose(ez[q], prodz.sum(axis=1), rtol=0, atol=1e-12 * scale), NAMES[q]
