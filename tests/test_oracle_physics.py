"""Closed forms and invariants of the substep, pinned on the oracle.

* laminar Poiseuille flow with a uniform heat source is an exact steady
  state: U and Theta unchanged, tau_w = 3 nu U_b, q_w = 3 (nu/Pr) Theta_b
  (wave00.f90:210-211, 621-622), mean pressure gradient sum(phi)/dt = 3 nu U_b
* the flow rate and the bulk scalar are held exactly (calculate_phi)
* one step keeps i kx u + dv/dy + i kz w = 0, v = dv/dy = 0 at the walls,
  Hermitian symmetry on kx = 0 and an empty kz Nyquist slot
* opposition control sets v(+-1) = -alpha v(y_d) exactly
"""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset, physical_wavenumbers
from channel_passive_amd.initial import make_state


def laminar(cfg, orc):
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    y = orc.y
    S[0, 0, 0] = 1.5 * cfg.bulk_u_value() * (1 - y * y)
    S[3, 0, 0] = 1.5 * cfg.bulk_t * (1 - y * y)
    return S


@pytest.mark.parametrize("name", ["re180_64", "tiny"])
def test_laminar_steady_state(oracle_lib, name):
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    S = laminar(cfg, orc)
    S0 = S.copy()
    RK = orc.init_stores(S)
    outs = orc.step(S, RK, cfg.dt)
    ub, nu = cfg.bulk_u_value(), cfg.nu
    assert np.abs(S[0, 0, 0] - S0[0, 0, 0]).max() <= 1e-12 * ub
    assert np.abs(S[3, 0, 0] - S0[3, 0, 0]).max() <= 1e-12 * cfg.bulk_t
    assert np.abs(S[1:3]).max() <= 1e-15 * ub  # FFT roundoff only
    assert abs(outs[-1]["tau_w"] - 3 * nu * ub) <= 1e-10 * 3 * nu * ub
    assert abs(outs[-1]["q_w"] - 3 * nu / cfg.pr * cfg.bulk_t) <= 1e-10 * 3 * nu / cfg.pr
    dpdx = sum(o["phi_u"] for o in outs) / cfg.dt
    assert abs(dpdx - 3 * nu * ub) <= 1e-8 * 3 * nu * ub


def divergence_and_walls(orc, cfg, S):
    from oracle import oracle as O
    _, kzi, kxp, kzp = physical_wavenumbers(cfg)
    worst = 0.0
    for j in range(cfg.nz):
        for k in range(cfg.nxh):
            if k == 0 and j == 0:
                continue
            u, v, w = S[0, j, k], S[1, j, k], S[2, j, k]
            dv = O.band_matvec(orc.d1, orc.bw, O.band_solve(orc.d0lu, orc.bw, v))
            scale = max(np.abs(dv).max(), np.abs(kxp[k] * u).max(), np.abs(kzp[j] * w).max(), 1e-300)
            div = 1j * kxp[k] * u + dv + 1j * kzp[j] * w
            worst = max(worst, np.abs(div).max() / scale, abs(dv[0]) / scale, abs(dv[-1]) / scale)
    return worst


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_step_invariants(oracle_lib, name):
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=21)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    assert divergence_and_walls(orc, cfg, S) <= 1e-11
    assert np.abs(S[1][..., 0]).max() == 0.0 and np.abs(S[1][..., -1]).max() == 0.0
    nh = cfg.nz // 2
    for j in range(1, nh):
        assert np.allclose(S[:, j, 0], np.conj(S[:, cfg.nz - j, 0]), rtol=0, atol=1e-13 * np.abs(S).max())
    assert not np.any(S[:, nh])
    from oracle import oracle as O
    cU = O.band_solve(orc.d0lu, orc.bw, S[0, 0, 0].copy())
    assert abs(0.5 * orc.integ @ cU.real - cfg.bulk_u_value()) <= 1e-13 * cfg.bulk_u_value()
    cT = O.band_solve(orc.d0lu, orc.bw, S[3, 0, 0].copy())
    assert abs(0.5 * orc.integ @ cT.real - cfg.bulk_t) <= 1e-13


def test_opposition_control_wall_values(oracle_lib):
    cfg = dataclasses.replace(preset("tiny"), opp_alpha=0.7, opp_yplus=25.0)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=4)
    RK = orc.init_stores(S)
    for s in (1, 2, 3):
        vd = orc.opp_planes(S)
        orc.substep(S, RK, s, cfg.dt)
        lo, hi = S[1][..., 0], S[1][..., -1]
        mask = np.ones(lo.shape, bool)
        mask[0, 0] = False  # the mean mode has no wall-normal velocity
        assert np.array_equal(lo[mask], (-cfg.opp_alpha * vd[0])[mask])
        assert np.array_equal(hi[mask], (-cfg.opp_alpha * vd[1])[mask])
    assert divergence_and_walls.__name__  # divergence still holds with moving walls:
    from oracle import oracle as O
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    for j in range(cfg.nz):
        for k in range(1, cfg.nxh):
            u, v, w = S[0, j, k], S[1, j, k], S[2, j, k]
            dv = O.band_matvec(orc.d1, orc.bw, O.band_solve(orc.d0lu, orc.bw, v))
            div = 1j * kxp[k] * u + dv + 1j * kzp[j] * w
            assert np.abs(div).max() <= 1e-11 * max(np.abs(dv).max(), 1e-300)


def test_opposition_detection_plane(oracle_lib):
    """v(y_d) is B(y_d) . D0^-1 v: for a polynomial profile of degree < order
    the interpolation is exact."""
    cfg = dataclasses.replace(preset("tiny"), opp_alpha=1.0, opp_yplus=18.0)
    orc = oracle_lib.Oracle(cfg)
    y = orc.y
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    S[1, 2, 3] = (1 - y * y) ** 2 * (0.3 + 0.1j)
    vd = orc.opp_planes(S)
    yd = 1.0 - cfg.opp_yplus / cfg.re_tau
    exact = (1 - yd * yd) ** 2 * (0.3 + 0.1j)
    assert abs(vd[0, 2, 3] - exact) <= 1e-14 and abs(vd[1, 2, 3] - exact) <= 1e-14


def test_detection_row_dense(oracle_lib):
    """The detection values r . v (r = LU(D0)^-T B(y_d), transposed band
    solve) equal B(y_d) . D0^-1 v computed with a dense solve, on random
    columns at both planes and for two band widths."""
    from oracle import oracle as O
    from tests.helpers import unpack_band
    for extra in ({}, {"ny": 65, "labs_k": 9}):
        cfg = dataclasses.replace(preset("tiny"), opp_alpha=1.0, opp_yplus=20.0, **extra)
        orc = oracle_lib.Oracle(cfg)
        rng = np.random.default_rng(6)
        S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
        S[1] = rng.standard_normal(S[1].shape) + 1j * rng.standard_normal(S[1].shape)
        vd = orc.opp_planes(S)
        D0 = unpack_band(orc.d0, orc.bw)
        C = np.linalg.solve(D0, S[1].reshape(-1, cfg.ny).T)  # columns: coefficients of each line
        yd = 1.0 - cfg.opp_yplus / cfg.re_tau
        for plane, y in ((0, -yd), (1, yd)):
            first, vals = O.bspline_row(cfg.ny, cfg.labs_k, cfg.knot_map, cfg.stretch, y)
            want = (vals[:, None] * C[first:first + len(vals)]).sum(axis=0).reshape(cfg.nz, cfg.nxh)
            scale = np.abs(S[1]).max()
            assert np.abs(vd[plane] - want).max() <= 1e-12 * scale, plane


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_mixed_scalar_laminar_closed_form(oracle_lib, name):
    """Uniform wall heat flux, mixed-type condition (source (nu/Pr) u/U_b,
    theta = 0 at the walls): on laminar Poiseuille flow U = 1.5 U_b (1 - y^2)
    the scalar Theta = (5 - 6 y^2 + y^4)/8 is steady, the wall flux per wall is
    q_w = nu/Pr, the bulk is free (phi_t = 0) and the mixed-mean Nusselt number
    (based on 2h) is 70/17."""
    from channel_passive_amd.config import PB_SCALAR_MIXED
    from channel_passive_amd.initial import scalar_mean_profile
    from channel_passive_amd import observables
    cfg = dataclasses.replace(preset(name), scalar_bc=PB_SCALAR_MIXED)
    orc = oracle_lib.Oracle(cfg)
    y = orc.y
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    S[0, 0, 0] = 1.5 * (1 - y * y)
    S[3, 0, 0] = scalar_mean_profile(cfg, y)
    assert np.allclose(S[3, 0, 0].real, (5 - 6 * y ** 2 + y ** 4) / 8, rtol=0, atol=1e-15)
    S0 = S.copy()
    RK = orc.init_stores(S)
    for _ in range(3):
        outs = orc.step(S, RK, cfg.dt)
    kappa = cfg.nu / cfg.pr
    assert np.abs(S[3, 0, 0] - S0[3, 0, 0]).max() <= 1e-12
    assert all(o["phi_t"] == 0.0 for o in outs)
    assert abs(outs[-1]["q_w"] - kappa) <= 1e-11 * kappa
    st = orc.statistics(S)
    quad = observables.Quadrature({"d0": orc.d0, "integ": orc.integ}, orc.bw)
    theta_m = observables.theta_mean(st, quad, "mixed")
    nu_mixed = observables.nusselt(outs[-1]["q_w"], cfg.nu, cfg.pr, theta_m)
    assert abs(nu_mixed - 70.0 / 17.0) <= 1e-11 * 70.0 / 17.0, nu_mixed


def test_mixed_scalar_perturbation_bounded(oracle_lib):
    """With the mixed-type condition the source acts on every mode: a
    perturbed state keeps theta = 0 at the walls on every line and finite."""
    from channel_passive_amd.config import PB_SCALAR_MIXED
    cfg = dataclasses.replace(preset("tiny"), scalar_bc=PB_SCALAR_MIXED)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=3)
    RK = orc.init_stores(S)
    for _ in range(5):
        orc.step(S, RK, cfg.dt)
    assert np.all(np.isfinite(S))
    assert np.abs(S[3][..., 0]).max() == 0.0 and np.abs(S[3][..., -1]).max() == 0.0
