"""Derived observables (channel_passive_amd/observables.py) on closed forms:
laminar Poiseuille flow with a uniform heat source has Cf = 12 / Re_b
(Re_b = 2 U_b h / nu), Nu = 6 with the plain bulk temperature and Nu = 5
with the velocity-weighted mixed mean; a single cosine mode of amplitude a
on top of a mean has rms a / sqrt(2)."""
import dataclasses

import numpy as np

from channel_passive_amd import observables as ob
from channel_passive_amd.config import preset


def laminar_run(oracle_lib, name="tiny"):
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    y = orc.y
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    ub, tb = cfg.bulk_u_value(), cfg.bulk_t
    S[0, 0, 0] = 1.5 * ub * (1 - y * y)
    S[3, 0, 0] = 1.5 * tb * (1 - y * y)
    RK = orc.init_stores(S)
    outs = orc.step(S, RK, cfg.dt)
    quad = ob.Quadrature({"d0": orc.d0, "integ": orc.integ}, orc.bw)
    return cfg, orc, S, outs[-1], quad


def test_band_to_dense_matches_quadrature(oracle_lib):
    """int (1 - y^2) dy = 4/3 and int y^4 dy = 2/5 through the packed D0."""
    cfg, orc, _, _, quad = laminar_run(oracle_lib)
    y = orc.y
    assert abs(quad(1 - y * y) - 4.0 / 3.0) <= 1e-13
    assert abs(quad(y ** 4) - 0.4) <= 1e-13


def test_laminar_cf_and_nu(oracle_lib):
    cfg, orc, S, out, quad = laminar_run(oracle_lib)
    ub, nu, tb = cfg.bulk_u_value(), cfg.nu, cfg.bulk_t
    re_b = 2 * ub / nu
    assert abs(ob.skin_friction(out["tau_w"], ub) - 12 / re_b) <= 1e-10 * 12 / re_b
    st = orc.statistics(S)
    tm_bulk = ob.theta_mean(st, quad, "bulk")
    tm_mixed = ob.theta_mean(st, quad, "mixed")
    assert abs(tm_bulk - tb) <= 1e-12 * tb
    assert abs(tm_mixed - 1.2 * tb) <= 1e-12 * tb
    assert abs(ob.nusselt(out["q_w"], nu, cfg.pr, tm_bulk) - 6.0) <= 1e-9
    assert abs(ob.nusselt(out["q_w"], nu, cfg.pr, tm_mixed) - 5.0) <= 1e-9


def test_rms_single_mode(oracle_lib):
    cfg = dataclasses.replace(preset("tiny"), u_gal=0.3)
    orc = oracle_lib.Oracle(cfg)
    y = orc.y
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    a = 1 - y * y
    S[0, 0, 0] = 2.0 * a           # mean U
    S[0, 0, 1] = 0.25 * a          # u' = 0.5 a cos(x): rms 0.5 a / sqrt(2)
    S[3, 1, 0] = 0.1 * a           # theta' on kz = +-1 (kx = 0 plane, with its conjugate)
    S[3, cfg.nz - 1, 0] = 0.1 * a  # theta' = 0.2 a cos(z): rms 0.2 a / sqrt(2)
    st = orc.statistics(S)
    r = ob.rms_profiles(st, u_gal=cfg.u_gal)
    assert np.allclose(r["u_rms"], 0.5 * a / np.sqrt(2), rtol=1e-12, atol=1e-14)
    assert np.allclose(r["t_rms"], 0.2 * a / np.sqrt(2), rtol=1e-12, atol=1e-14)
    assert not np.any(r["v_rms"]) and not np.any(r["w_rms"])
