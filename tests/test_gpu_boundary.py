"""GPU: the per-routine entry points of the C ABI (y_to_x, dft_x_realspace,
x_to_y + rescale_data, get_pressure, LABS_lu / LABS_bs / LABS_mat_mul) on
host arrays in the reference layouts.

* y_to_x and x_to_y against numpy.fft (an independent library) with the
  reference's 3/2 padding and truncation positions (module_ppdfftw.f90:
  810-817, 846-854): <= 1e-13 relative L2 per field (FFT rounding).
* y_to_x -> dft_x_realspace -> x_to_y(scale = rescale) is bit-identical to
  the fused pb_nonlinear_host, which test_gpu_parity.py holds against the
  oracle; the x-pass velmax equals the oracle's.
* get_pressure on the oracle's own products: <vv> bitwise, p to 1e-13.
* the LABS operations bit-identical to the oracle's restatement of
  module_labs.f90 (bands 3, 4, 5; stacks of matrices and columns).
"""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state
from tests.helpers import rel

pytestmark = pytest.mark.gpu

FFT_TOL = 1e-13


@pytest.fixture(scope="module")
def capi():
    from channel_passive_amd import capi as C
    return C


def cfgs():
    return {"re180_64": preset("re180_64"),
            "flat": dataclasses.replace(preset("re1000"), ny=17, name="flat"),  # production x/z sizes
            "tiny_opp": dataclasses.replace(preset("tiny"), opp_alpha=0.5, opp_yplus=20.0)}


CFGS = cfgs()


def pad_z(cfg, S):
    """(f, nz, nxh, ny) spectral -> (f, nz32, nxh, ny) with kz >= 0 in front,
    negative kz at the tail, the Nyquist slot and the gap zero."""
    nh, nz32 = cfg.nz // 2, 3 * cfg.nz // 2
    Z = np.zeros((S.shape[0], nz32) + S.shape[2:], complex)
    Z[:, :nh] = S[:, :nh]
    Z[:, nz32 - (nh - 1):] = S[:, nh + 1:]
    return Z


@pytest.mark.parametrize("name", ["re180_64", "flat"])
def test_y_to_x_against_numpy(capi, name):
    cfg = CFGS[name]
    S = make_state(cfg, seed=31)
    ch = capi.Channel(cfg)
    got = ch.y_to_x_host(S)
    nz32 = 3 * cfg.nz // 2
    want = np.fft.ifft(pad_z(cfg, S), axis=1) * nz32          # (4, nz32, nxh, ny)
    want = np.transpose(want, (0, 3, 1, 2))                  # (4, ny, nz32, nxh)
    for f in range(4):
        e = rel(got[f], want[f])
        print(f"{name} field {f}: {e:.2e}")
        assert e <= FFT_TOL, (f, e)
    ch.close()


@pytest.mark.parametrize("name,nf,scale", [("re180_64", 9, 1.0), ("re180_64", 3, 0.25), ("flat", 8, None)])
def test_x_to_y_against_numpy(capi, name, nf, scale):
    cfg = CFGS[name]
    nh, nz32, nx32 = cfg.nz // 2, 3 * cfg.nz // 2, 3 * cfg.nx // 2
    scale = 1.0 / (nx32 * nz32) if scale is None else scale
    rng = np.random.default_rng(nf)
    ch = capi.Channel(cfg)
    shp = ch.pshape(nf)
    Up = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
    got = ch.x_to_y_host(Up, scale)
    F = np.fft.fft(np.transpose(Up, (0, 2, 3, 1)), axis=1)   # (nf, nz32, nxh, ny)
    want = np.zeros((nf, cfg.nz, cfg.nxh, cfg.ny), complex)
    want[:, :nh] = F[:, :nh]
    want[:, nh + 1:] = F[:, nz32 - (nh - 1):]
    want *= scale
    assert not np.any(got[:, nh])
    for f in range(nf):
        e = rel(got[f], want[f])
        assert e <= FFT_TOL, (f, e)
    ch.close()


@pytest.mark.parametrize("name", ["re180_64", "flat", "tiny_opp"])
def test_stage_chain_equals_fused(capi, oracle_lib, name):
    cfg = CFGS[name]
    S = make_state(cfg, seed=5)
    ch = capi.Channel(cfg)
    P_fused, vm_fused = ch.nonlinear_host(S, want_vv=True)
    Up = ch.y_to_x_host(S)
    Q, vm = ch.realspace_host(Up, want_vv=True)
    assert np.array_equal(Q[9], Up[1])  # the reference keeps v in slot 10
    nx32, nz32 = 3 * cfg.nx // 2, 3 * cfg.nz // 2
    P = ch.x_to_y_host(np.ascontiguousarray(Q[:9]), 1.0 / (nx32 * nz32))
    for slot in range(9):
        assert np.array_equal(P[slot], P_fused[slot]), slot
    assert vm == vm_fused
    _, vm_orc = oracle_lib.Oracle(cfg).nonlinear(S)
    assert abs(vm - vm_orc) <= 1e-13 * vm_orc
    # without v*v: slot 5 comes back zero, the others unchanged
    Q2, vm2 = ch.realspace_host(Up, want_vv=False)
    assert not np.any(Q2[5]) and abs(vm2 - vm) <= 1e-15 * vm  # the streamed x pass may pair FMAs differently
    ch.close()


@pytest.mark.parametrize("name", ["re180_64", "tiny_opp"])
def test_pressure_host(capi, oracle_lib, name):
    cfg = CFGS[name]
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=8)
    P, _ = orc.nonlinear(S, want_vv=True)
    Pr = P.copy()
    vv_ref = orc.pressure(S, Pr)
    ch = capi.Channel(cfg)
    p, vv = ch.pressure_host(S, np.ascontiguousarray(P))
    assert np.array_equal(vv, vv_ref)
    e = rel(p, Pr[5])
    print(f"{name}: p rel L2 {e:.2e}")
    assert e <= 1e-13
    ch.close()


def random_band(oracle_lib, n, bw, rng):
    from oracle import oracle as O
    F = np.zeros((n, n))
    for r in range(n):
        lo, hi = max(0, r - bw), min(n, r + bw + 1)
        F[r, lo:hi] = rng.standard_normal(hi - lo)
        F[r, r] += 2.0 * bw + 1.0
    # the near-boundary rows reach 2bw columns in, as the B-spline bands do
    F[0, bw + 1:2 * bw + 1] = rng.standard_normal(bw)
    F[n - 1, n - 1 - 2 * bw:n - 1 - bw] = rng.standard_normal(bw)
    return O.full_to_band(F, bw)


@pytest.mark.parametrize("bw", [3, 4, 5])
def test_labs_bitwise(capi, oracle_lib, bw):
    from oracle import oracle as O
    rng = np.random.default_rng(bw)
    n = 37
    A = np.stack([random_band(oracle_lib, n, bw, rng) for _ in range(5)])
    LU = capi.labs_lu(A, bw)
    for m in range(5):
        assert np.array_equal(LU[m], O.band_lu(A[m], bw)), m
    B = rng.standard_normal((7, n)) + 1j * rng.standard_normal((7, n))
    X = capi.labs_bs(LU[2], bw, B)
    Y = capi.labs_mat_mul(A[3], bw, B)
    for q in range(7):
        assert np.array_equal(X[q], O.band_solve(LU[2], bw, B[q])), q
        assert np.array_equal(Y[q], O.band_matvec(A[3], bw, B[q])), q
    # a single call is the reference's LABS_bs / LABS_mat_mul
    assert np.array_equal(capi.labs_bs(LU[0], bw, B[0]), O.band_solve(LU[0], bw, B[0]))


def test_labs_rejects_bad_sizes(capi):
    with pytest.raises(capi.PbError, match="labs"):
        capi.labs_lu(np.zeros((5, 9)), 4)  # n = 5 < 2 bw + 1
