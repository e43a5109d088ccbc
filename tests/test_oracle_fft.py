"""Pins of the oracle's transforms: the DFT against numpy.fft, and the whole
nonlinear stage (y_to_x, dft_x_realspace, x_to_y, rescale_data) against an
independent numpy restatement of the reference's FFTW calls
(module_ppdfftw.f90:794-861, dft_x_realspace.f90:68-213, rescale_data.f90)."""
import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state
from tests.helpers import rel


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8, 12, 96, 144, 240, 288, 1152, 1536, 7 * 11])
def test_dft_matches_numpy(oracle_lib, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    f = oracle_lib.fft(x, -1)
    b = oracle_lib.fft(x, +1)
    assert np.abs(f - np.fft.fft(x)).max() <= 1e-14 * np.abs(f).max()
    assert np.abs(b - np.fft.ifft(x) * n).max() <= 1e-14 * np.abs(b).max()


def numpy_nonlinear(cfg, S):
    """numpy restatement: zero-pad kz to 3nz/2 (negative kz at the tail, the
    Nyquist slot dropped), unnormalised backward c2c along z, c2r along x from
    nx/2 modes zero padded to 3nx/4+1 (Im of the DC bin ignored), products,
    r2c, keep nx/2 modes, forward c2c along z, keep nz/2 + tail, Nyquist = 0,
    times 1/(nx32*nz32)."""
    nxh, nz, ny = cfg.nxh, cfg.nz, cfg.ny
    nh, nz32, nx32 = nz // 2, 3 * nz // 2, 3 * cfg.nx // 2
    pad = np.zeros((4, nz32, nxh, ny), complex)
    pad[:, :nh] = S[:, :nh]
    pad[:, 2 * nh + 1:] = S[:, nh + 1:]
    G = np.fft.ifft(pad, axis=1) * nz32                       # (4, z, kx, y)
    X = np.zeros((4, nz32, nx32 // 2 + 1, ny), complex)
    X[:, :, :nxh] = G
    X[:, :, 0] = X[:, :, 0].real
    phys = np.fft.irfft(X, n=nx32, axis=2) * nx32             # (4, z, x, y)
    u, v, w, t = phys
    prods = [u * u - v * v, u * v, w * w - v * v, v * w, u * w, v * v, u * t, v * t, w * t, v]
    P = np.zeros((10, nz, nxh, ny), complex)
    for f, p in enumerate(prods):
        H = np.fft.rfft(p, axis=1)[:, :nxh]                   # (z, kx, y)
        Z = np.fft.fft(H, axis=0)
        P[f, :nh] = Z[:nh]
        P[f, nh + 1:] = Z[2 * nh + 1:]
    return P / (nx32 * nz32)


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_nonlinear_stage_vs_numpy(oracle_lib, name):
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=9)
    S[:, 0, 0] += 0.01j  # an imaginary DC part that c2r must ignore
    P, vmax = orc.nonlinear(S, want_vv=True)
    Pn = numpy_nonlinear(cfg, S)
    for f in range(10):
        assert rel(P[f], Pn[f]) <= 1e-13, f


def test_dealiasing_exact(oracle_lib):
    """3/2 rule: the product of two resolved modes is computed without
    aliasing, e.g. u = cos(3x) cos(2z) gives u^2 with the exact modes."""
    cfg = preset("tiny")
    orc = oracle_lib.Oracle(cfg)
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    kx, kz = cfg.nxh - 1, cfg.nz // 2 - 1  # highest resolved modes
    S[0, kz, kx] = 0.25
    S[0, cfg.nz - kz, kx] = 0.25
    P, _ = orc.nonlinear(S)
    # u = cos(kx x)cos(kz z): u^2 has modes (0,0), (0,+-2kz), (2kx,...) which are all
    # outside the resolved set except (0,0), (0,+-2kz) -> (0,0) = 1/4 only
    uu = P[0]
    assert abs(uu[0, 0, 0] - 0.25) <= 1e-15
    mask = np.ones(uu.shape[:2], bool)
    mask[0, 0] = False
    assert np.abs(uu[mask]).max() <= 1e-15
