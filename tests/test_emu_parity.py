"""The product's per-system solver (channel_passive_amd/csrc/solver.h) built
for the CPU equals the oracle bit for bit, on every substep, with and
without opposition control.  This is the same source the GPU kernel runs."""
import ctypes
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state

dp = ctypes.POINTER(ctypes.c_double)


@pytest.mark.parametrize("name,alpha,sbc", [("tiny", 0.0, 0), ("tiny", 0.8, 0), ("re180_64", 0.0, 0),
                                            ("tiny", 0.8, 1)])
def test_wavespace_bitwise_cpu(oracle_lib, emu_lib, name, alpha, sbc):
    cfg = dataclasses.replace(preset(name), opp_alpha=alpha, opp_yplus=20.0, scalar_bc=sbc)
    orc = oracle_lib.Oracle(cfg)
    c = cfg.to_c()
    S = make_state(cfg, seed=13)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    for s in (1, 2, 3):
        P, _ = orc.nonlinear(S)
        vd = orc.opp_planes(S) if alpha else None
        Ud = np.ascontiguousarray(np.transpose(P[:9], (0, 3, 1, 2)))
        RKd = np.ascontiguousarray(np.transpose(RK, (0, 3, 1, 2)))
        us = np.ascontiguousarray(np.transpose(S[0], (2, 0, 1)))  # u of the input state, device layout
        m_e = np.zeros(4)
        assert emu_lib.emu_wavespace(ctypes.byref(c), s, ctypes.c_double(cfg.dt), Ud.ctypes.data_as(dp),
                                     RKd.ctypes.data_as(dp), None if vd is None else vd.ctypes.data_as(dp),
                                     us.ctypes.data_as(dp), m_e.ctypes.data_as(dp)) == 0
        m_o = orc.wavespace(s, cfg.dt, P, RK, vd, usrc=S[0])
        Ue, RKe = np.transpose(Ud, (0, 2, 3, 1)), np.transpose(RKd, (0, 2, 3, 1))
        for f in (0, 1, 2, 6):
            assert np.array_equal(Ue[f], P[f]), (s, f)
        for i in range(3):
            assert np.array_equal(RKe[i], RK[i]), (s, i)
        assert np.array_equal(m_e, m_o)
        for i, f in enumerate((0, 1, 2, 6)):
            S[i] = P[f]


@pytest.mark.parametrize("name,alpha,extra", [("tiny", 0.0, {}), ("tiny", 0.8, {}), ("re180_64", 0.0, {}),
                                              ("tiny", 0.5, {"knot_map": 0, "stretch": 0.0, "labs_k": 6, "ny": 34}),
                                              ("tiny", 0.0, {"ny": 65, "labs_k": 9}),
                                              ("tiny", 0.8, {"scalar_bc": 1})])
def test_fused_wavespace_bitwise_cpu(oracle_lib, emu_lib, name, alpha, extra):
    """The seven-sweep organisation of the device solver (solver_fused.h)
    reproduces the oracle bit for bit (bw = 4, 3 and 5 grids; the mixed-type
    scalar condition)."""
    cfg = dataclasses.replace(preset(name), opp_alpha=alpha, opp_yplus=20.0, **extra)
    orc = oracle_lib.Oracle(cfg)
    assert orc.bw in (3, 4, 5)
    c = cfg.to_c()
    S = make_state(cfg, seed=17)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    for s in (1, 2, 3):
        P, _ = orc.nonlinear(S)
        vd = orc.opp_planes(S) if alpha else None
        Ud = np.ascontiguousarray(np.transpose(P[:9], (0, 3, 1, 2)))
        RKd = np.ascontiguousarray(np.transpose(RK, (0, 3, 1, 2)))
        us = np.ascontiguousarray(np.transpose(S[0], (2, 0, 1)))
        m_e = np.zeros(4)
        assert emu_lib.emu_wavespace_fused(ctypes.byref(c), s, ctypes.c_double(cfg.dt), Ud.ctypes.data_as(dp),
                                           RKd.ctypes.data_as(dp), None if vd is None else vd.ctypes.data_as(dp),
                                           us.ctypes.data_as(dp), m_e.ctypes.data_as(dp)) == 0
        m_o = orc.wavespace(s, cfg.dt, P, RK, vd, usrc=S[0])
        Ue, RKe = np.transpose(Ud, (0, 2, 3, 1)), np.transpose(RKd, (0, 2, 3, 1))
        for f in (0, 1, 2, 6):
            assert np.array_equal(Ue[f], P[f]), (s, f, np.abs(Ue[f] - P[f]).max())
        for i in range(3):
            assert np.array_equal(RKe[i], RK[i]), (s, i, np.abs(RKe[i] - RK[i]).max())
        assert np.array_equal(m_e, m_o)
        for i, f in enumerate((0, 1, 2, 6)):
            S[i] = P[f]


@pytest.mark.parametrize("name,extra", [("tiny", {}), ("re180_64", {}),
                                        ("tiny", {"knot_map": 0, "stretch": 0.0, "labs_k": 6, "ny": 34}),
                                        ("tiny", {"ny": 65, "labs_k": 9})])
def test_opp_planes_bitwise_cpu(oracle_lib, emu_lib, name, extra):
    """The per-system detection dot products (solver.h opp_dots over the rows
    LU(D0)^-T B(y_d) of bspline.cpp detection_row) give the oracle's v(y_d)
    bit for bit."""
    cfg = dataclasses.replace(preset(name), opp_alpha=0.7, opp_yplus=20.0, **extra)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=29)
    vd = orc.opp_planes(S)
    V = np.ascontiguousarray(np.transpose(S[1], (2, 0, 1)))  # (nz, nxh, ny) -> [y][kz][kx]
    out = np.zeros((2, cfg.nz, cfg.nxh), np.complex128)
    c = cfg.to_c()
    assert emu_lib.emu_opp_planes(ctypes.byref(c), V.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    assert np.array_equal(out, vd), np.abs(out - vd).max()


@pytest.mark.parametrize("name", ["tiny", "re180_64", "re180_192"])
def test_pressure_bitwise_cpu(oracle_lib, emu_lib, name):
    """The product's pressure code (solver.h ppe_wall_values, solve_ppe,
    mean_vv_row) in the device layout equals the oracle bit for bit."""
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=31)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    P, _ = orc.nonlinear(S, want_vv=True)
    Ud = np.ascontiguousarray(np.transpose(P[:9], (0, 3, 1, 2)))
    V = np.ascontiguousarray(np.transpose(S[1], (2, 0, 1)))
    vv_o = orc.pressure(S, P)
    vv_e = np.zeros(cfg.ny)
    c = cfg.to_c()
    assert emu_lib.emu_pressure(ctypes.byref(c), V.ctypes.data_as(dp), Ud.ctypes.data_as(dp),
                                vv_e.ctypes.data_as(dp)) == 0
    assert np.array_equal(vv_e, vv_o)
    pe = np.transpose(Ud[5], (1, 2, 0))
    assert np.array_equal(pe, P[5]), np.abs(pe - P[5]).max()


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_statistics_bitwise_cpu(oracle_lib, emu_lib, name):
    """The product's statistics code (solver.h mean_uv_row, mean_line_stats)
    in the device layout equals the oracle bit for bit; the centre values
    (from the product's own basis row at y = 0) to 1e-15."""
    cfg = dataclasses.replace(preset(name), u_gal=0.1)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=41)
    ref = orc.statistics(S)
    U = np.zeros((9, cfg.ny, cfg.nz, cfg.nxh), np.complex128)
    for i, f in enumerate((0, 1, 2, 6)):
        U[f] = np.transpose(S[i], (2, 0, 1))
    out = np.zeros(16 * cfg.ny + 2)
    c = cfg.to_c()
    assert emu_lib.emu_statistics(ctypes.byref(c), U.ctypes.data_as(dp), out.ctypes.data_as(dp)) == 0
    n = cfg.ny
    for i, k in enumerate(orc.STAT_NAMES):
        assert np.array_equal(out[i * n:(i + 1) * n], ref[k]), k
    assert abs(out[16 * n] - ref["u_center"]) <= 1e-15 * abs(ref["u_center"])
    assert abs(out[16 * n + 1] - ref["t_center"]) <= 1e-15 * max(abs(ref["t_center"]), 1.0)


@pytest.mark.parametrize("name,alpha,extra", [("tiny", 0.0, {}), ("tiny", 0.8, {}), ("tiny", 0.6, {"scalar_bc": 1}),
                                              ("re180_64", 0.5, {})])
def test_actuated_wavespace_bitwise_cpu(oracle_lib, emu_lib, name, alpha, extra):
    """Wall actuation: v(wall) = -alpha v(y_d) + a(kx, kz) on both walls.  The
    device solver (solver_fused.h) with an actuation field equals the oracle
    with the same field (orc_set_control) bit for bit; the field changes the
    result, and a zero field gives the unactuated solve."""
    cfg = dataclasses.replace(preset(name), opp_alpha=alpha, opp_yplus=15.0, **extra)
    orc = oracle_lib.Oracle(cfg)
    c = cfg.to_c()
    rng = np.random.default_rng(3)
    act = 1e-2 * (rng.standard_normal((cfg.nz, cfg.nxh)) + 1j * rng.standard_normal((cfg.nz, cfg.nxh)))
    S = make_state(cfg, seed=19)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    orc.set_control(alpha, act)
    for s in (1, 2, 3):
        P, _ = orc.nonlinear(S)
        vd = orc.opp_planes(S) if alpha else None
        Ud = np.ascontiguousarray(np.transpose(P[:9], (0, 3, 1, 2)))
        RKd = np.ascontiguousarray(np.transpose(RK, (0, 3, 1, 2)))
        us = np.ascontiguousarray(np.transpose(S[0], (2, 0, 1)))
        Ud0, RKd0 = Ud.copy(), RKd.copy()
        m_e, m_0 = np.zeros(4), np.zeros(4)
        args = (None if vd is None else vd.ctypes.data_as(dp), us.ctypes.data_as(dp))
        assert emu_lib.emu_wavespace_fused_act(ctypes.byref(c), s, ctypes.c_double(cfg.dt), Ud.ctypes.data_as(dp),
                                               RKd.ctypes.data_as(dp), *args, act.ctypes.data_as(dp),
                                               m_e.ctypes.data_as(dp)) == 0
        assert emu_lib.emu_wavespace_fused_act(ctypes.byref(c), s, ctypes.c_double(cfg.dt), Ud0.ctypes.data_as(dp),
                                               RKd0.ctypes.data_as(dp), *args, None, m_0.ctypes.data_as(dp)) == 0
        assert not np.array_equal(Ud0[1], Ud[1])
        m_o = orc.wavespace(s, cfg.dt, P, RK, vd, usrc=S[0])
        Ue, RKe = np.transpose(Ud, (0, 2, 3, 1)), np.transpose(RKd, (0, 2, 3, 1))
        for f in (0, 1, 2, 6):
            assert np.array_equal(Ue[f], P[f]), (s, f, np.abs(Ue[f] - P[f]).max())
        for i in range(3):
            assert np.array_equal(RKe[i], RK[i]), (s, i)
        assert np.array_equal(m_e, m_o)
        for i, f in enumerate((0, 1, 2, 6)):
            S[i] = P[f]
