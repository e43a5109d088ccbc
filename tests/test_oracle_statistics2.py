"""get_statistics_2 (get_statistics.f90:376-539; the moments of
dft_x_realspace.f90:128-484) on the oracle, pinned against an independent
physical-space restatement with numpy.fft: the state and the oracle's
pressure are transformed to the 3/2-padded physical grid, the moments are
plain averages of powers there, the correlations averages of products
(Parseval makes them equal to the reference's spectral sums: the padded
grid resolves every quadratic product), x and z derivatives are spectral,
y derivatives D1 D0^-1 on each line (dense)."""
import numpy as np
import pytest

from channel_passive_amd.config import physical_wavenumbers, preset
from channel_passive_amd.initial import make_state
from tests.helpers import unpack_band


def to_physical(cfg, F):
    """(nz, nxh, ny) spectral -> (ny, nz32, nx32) physical values
    (unnormalised inverse transforms, FFTW's c2r convention in x)."""
    nh, nz32, nx32 = cfg.nz // 2, 3 * cfg.nz // 2, 3 * cfg.nx // 2
    Z = np.zeros((nz32, cfg.nxh, cfg.ny), complex)
    Z[:nh] = F[:nh]
    Z[nz32 - (nh - 1):] = F[nh + 1:]
    G = np.fft.ifft(Z, axis=0) * nz32                       # (nz32, nxh, ny)
    X = np.zeros((nz32, nx32 // 2 + 1, cfg.ny), complex)
    X[:, :cfg.nxh] = G
    X[:, 0].imag = 0.0                                      # c2r ignores Im of the DC bin
    phys = np.fft.irfft(X, n=nx32, axis=1) * nx32           # (nz32, nx32, ny)
    return np.transpose(phys, (2, 0, 1))


def numpy_stats2(cfg, orc, S, p):
    n = cfg.ny
    D0, D1 = unpack_band(orc.d0, orc.bw), unpack_band(orc.d1, orc.bw)
    dy = D1 @ np.linalg.inv(D0)
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    u, v, w = (to_physical(cfg, S[f]) for f in range(3))
    pp = to_physical(cfg, p)
    mean = lambda a: a.mean(axis=(1, 2))
    out = {}
    for name, val in (("uuu", u ** 3), ("vvv", v ** 3), ("www", w ** 3), ("uuuu", u ** 4), ("vvvv", v ** 4),
                      ("wwww", w ** 4), ("u6", u ** 6), ("v6", v ** 6), ("w6", w ** 6), ("u8", u ** 8),
                      ("v8", v ** 8), ("w8", w ** 8), ("u10", u ** 10), ("v10", v ** 10), ("w10", w ** 10)):
        m = mean(val)
        m[0] = m[-1] = 0.0
        out[name] = m
    out["-duuvdy"] = dy @ mean(-u * u * v)
    out["-dvvvdy"] = dy @ mean(-v ** 3)
    out["-dvwwdy"] = dy @ mean(-v * w * w)
    out["-duvvdy"] = dy @ mean(-u * v * v)
    out["up"], out["vp"], out["wp"], out["pp"] = mean(u * pp), mean(v * pp), mean(w * pp), mean(pp * pp)
    out["-dpudy"] = -(dy @ out["up"])
    out["-2dpvdy"] = -2.0 * (dy @ out["vp"])
    dudx = to_physical(cfg, 1j * kxp[None, :, None] * S[0])
    dwdz = to_physical(cfg, 1j * kzp[:, None, None] * S[2])
    dvdx = to_physical(cfg, 1j * kxp[None, :, None] * S[1])
    dudy = to_physical(cfg, np.einsum("ij,abj->abi", dy, S[0]))
    dvdy = to_physical(cfg, np.einsum("ij,abj->abi", dy, S[1]))
    out["2pdudx"] = mean(2 * pp * dudx)
    out["2pdvdy"] = mean(2 * pp * dvdy)
    out["2pdwdz"] = mean(2 * pp * dwdz)
    out["pdudy+pdvdx"] = mean(pp * dudy) + mean(pp * dvdx)
    return out


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_statistics2_against_physical_space(oracle_lib, name):
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=23)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    st, P, _ = orc.statistics2(S)
    ref = numpy_stats2(cfg, orc, S, P[5])
    for k in orc.STAT2_NAMES:
        scale = max(np.abs(ref[k]).max(), 1e-300)
        e = np.abs(st[k] - ref[k]).max() / scale
        assert e <= 1e-10, (k, e)


def test_statistics2_pin_fails_on_planted_mistake(oracle_lib):
    from oracle import oracle as O
    cfg = preset("tiny")
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=23)
    with O.mutant(10):
        st, P, _ = orc.statistics2(S)
    ref = numpy_stats2(cfg, orc, S, P[5])
    e = np.abs(st["2pdudx"] - ref["2pdudx"]).max() / np.abs(ref["2pdudx"]).max()
    assert e > 1e-3
