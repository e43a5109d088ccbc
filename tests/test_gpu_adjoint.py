"""GPU: the discrete adjoint of the RK3-CN substep (the transposed
wall-normal stage, the transposed nonlinear stage, the opposition control
and the wall actuation) and the checkpointed gradient driver.

* The transposed wall-normal stage equals its CPU build (the same source,
  solver_adj.h, whose dot-product identity with the forward is tested on
  the CPU).
* The transposed nonlinear stage and the transposed substep satisfy the
  dot-product identity with the GPU forward: the nonlinear stage is
  quadratic and the rest affine, so (F(x + d) - F(x - d)) / 2 is exactly
  the linearisation at x.
* The gradient of J = mean over the window of tau_w + 0.5 q_w with respect
  to the opposition gain and the actuation field agrees with central finite
  differences of the GPU forward to 1e-6, with those of the oracle's
  forward to 1e-5, and through a torch.autograd.Function harness as well.
"""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state

pytestmark = pytest.mark.gpu

PROD = (0, 1, 2, 3, 4, 6, 7, 8)


@pytest.fixture(scope="module")
def capi():
    from channel_passive_amd import capi as C
    return C


def rdot(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def cfg_of(name, alpha, sbc, yplus=15.0):
    return dataclasses.replace(preset(name), opp_alpha=alpha, opp_yplus=yplus, scalar_bc=sbc)


@pytest.mark.parametrize("name,alpha,sbc", [("tiny", 0.8, 0), ("tiny", 0.6, 1), ("re180_64", 0.5, 0)])
def test_adj_wavespace_matches_cpu_build(capi, emu_lib, name, alpha, sbc):
    import ctypes
    cfg = cfg_of(name, alpha, sbc)
    ch = capi.Channel(cfg)
    rng = np.random.default_rng(7)
    dp = ctypes.POINTER(ctypes.c_double)
    for s in (1, 2, 3):
        AU = crand(rng, 9, cfg.nz, cfg.nxh, cfg.ny)
        ARK = crand(rng, 3, cfg.nz, cfg.nxh, cfg.ny)
        vd = crand(rng, 2, cfg.nz, cfg.nxh)
        tb, qb = (0.7, -0.3) if s == 3 else (0.0, 0.0)
        AUg, ARKg = AU.copy(), ARK.copy()
        avd, act, al = ch.adj_wavespace_host(s, cfg.dt, AUg, ARKg, vd, tb, qb)
        dev = lambda a: np.ascontiguousarray(np.transpose(a, (0, 3, 1, 2)))
        AUe, ARKe = dev(AU), dev(ARK)
        avd_e = np.zeros(2 * cfg.nz * cfg.nxh, complex)
        act_e = np.zeros(cfg.nz * cfg.nxh, complex)
        al_e = np.zeros(cfg.nz * cfg.nxh)
        c = cfg.to_c()
        assert emu_lib.emu_adj_wavespace(ctypes.byref(c), s, ctypes.c_double(cfg.dt), AUe.ctypes.data_as(dp),
                                         ARKe.ctypes.data_as(dp), np.ascontiguousarray(vd).ctypes.data_as(dp),
                                         avd_e.ctypes.data_as(dp), act_e.ctypes.data_as(dp),
                                         al_e.ctypes.data_as(dp), ctypes.c_double(tb), ctypes.c_double(qb)) == 0
        ref = lambda a: np.transpose(a, (0, 2, 3, 1))
        for f in (*PROD, 5):
            e = np.abs(AUg[f] - ref(AUe)[f]).max() / max(np.abs(ref(AUe)[f]).max(), 1e-300)
            assert e <= 1e-13, (s, f, e)
        assert np.abs(ARKg - ref(ARKe)).max() <= 1e-13 * np.abs(ARKe).max()
        assert np.abs(avd.ravel() - avd_e).max() <= 1e-13 * max(np.abs(avd_e).max(), 1e-300)
        assert np.abs(act.ravel() - act_e).max() <= 1e-13 * max(np.abs(act_e).max(), 1e-300)
        assert np.abs(al.ravel() - al_e).max() <= 1e-13 * max(np.abs(al_e).max(), 1e-300)
    ch.close()


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_adj_nonlinear_dot(capi, name):
    cfg = preset(name)
    ch = capi.Channel(cfg)
    S = make_state(cfg, seed=11)
    rng = np.random.default_rng(3)
    d = 1e-2 * crand(rng, *S.shape)
    Pp, _ = ch.nonlinear_host(S + d)
    Pm, _ = ch.nonlinear_host(S - d)
    dN = (Pp - Pm) / 2
    Pb = crand(rng, 10, cfg.nz, cfg.nxh, cfg.ny)
    aS = ch.adj_nonlinear_host(S, Pb)
    lhs = sum(rdot(dN[f], Pb[f]) for f in PROD)
    rhs = rdot(d, aS)
    scale = np.sqrt(sum(rdot(dN[f], dN[f]) for f in PROD) * sum(rdot(Pb[f], Pb[f]) for f in PROD))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, abs(lhs - rhs) / scale)
    ch.close()


def _substep(ch, S, RK, s, dt):
    ch.set_state(np.ascontiguousarray(S), np.ascontiguousarray(RK))
    out = ch.substep(s, dt)
    S1, RK1 = ch.get_state()
    return S1, RK1, out


@pytest.mark.parametrize("name,alpha,sbc", [("tiny", 0.8, 0), ("tiny", 0.6, 1), ("re180_64", 0.5, 1)])
def test_adj_substep_dot(capi, name, alpha, sbc):
    cfg = cfg_of(name, alpha, sbc)
    ch = capi.Channel(cfg)
    S = make_state(cfg, seed=5)
    ch.set_state(S)
    ch.init_stores()
    _, RK = ch.get_state()
    rng = np.random.default_rng(17)
    act = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
    ch.set_control(alpha, act)
    for s in (1, 2, 3):
        d = 1e-2 * crand(rng, *S.shape)
        r = 1e-2 * crand(rng, *RK.shape)
        Sp, Rp, op = _substep(ch, S + d, RK + r, s, cfg.dt)
        Sm, Rm, om = _substep(ch, S - d, RK - r, s, cfg.dt)
        dS, dR = (Sp - Sm) / 2, (Rp - Rm) / 2
        yS, yR = crand(rng, *S.shape), crand(rng, *RK.shape)
        tb, qb = (0.9, 0.4) if s == 3 else (0.0, 0.0)
        lhs = rdot(dS, yS) + rdot(dR, yR)
        if s == 3:
            lhs += (op["tau_w"] - om["tau_w"]) / 2 * tb + (op["q_w"] - om["q_w"]) / 2 * qb
        aS, aR = yS.copy(), yR.copy()
        ab, act_b = ch.adj_substep(s, cfg.dt, S, aS, aR, tb, qb)
        rhs = rdot(d, aS) + rdot(r, aR)
        scale = np.sqrt((rdot(dS, dS) + rdot(dR, dR)) * (rdot(yS, yS) + rdot(yR, yR)))
        assert abs(lhs - rhs) <= 1e-11 * scale, (s, lhs, rhs, abs(lhs - rhs) / scale)
        # the controls: the substep is affine in alpha and in the actuation
        h = 0.25
        ch.set_control(alpha + h, act)
        Sa, Ra, oa = _substep(ch, S, RK, s, cfg.dt)
        ch.set_control(alpha - h, act)
        Sb, Rb, ob = _substep(ch, S, RK, s, cfg.dt)
        ch.set_control(alpha, act)
        dl = (rdot(Sa - Sb, yS) + rdot(Ra - Rb, yR)) / (2 * h)
        if s == 3:
            dl += ((oa["tau_w"] - ob["tau_w"]) * tb + (oa["q_w"] - ob["q_w"]) * qb) / (2 * h)
        assert abs(dl - ab) <= 1e-9 * (abs(dl) + abs(ab) + 1e-300), (s, dl, ab)
        e = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
        ch.set_control(alpha, act + e)
        Sa, Ra, oa = _substep(ch, S, RK, s, cfg.dt)
        ch.set_control(alpha, act - e)
        Sb, Rb, ob = _substep(ch, S, RK, s, cfg.dt)
        ch.set_control(alpha, act)
        da = (rdot(Sa - Sb, yS) + rdot(Ra - Rb, yR)) / 2
        if s == 3:
            da += ((oa["tau_w"] - ob["tau_w"]) * tb + (oa["q_w"] - ob["q_w"]) * qb) / 2
        assert abs(da - rdot(e, act_b)) <= 1e-8 * (abs(da) + 1e-300), (s, da, rdot(e, act_b))
    ch.close()


def _J(ch, S0, RK0, alpha, act, nsteps, dt, wq=0.5):
    ch.set_control(alpha, act)
    ch.set_state(S0, RK0)
    j = 0.0
    for _ in range(nsteps):
        o = ch.step(dt)
        j += o[-1]["tau_w"] + wq * o[-1]["q_w"]
    return j / nsteps


@pytest.mark.parametrize("sbc", [0, 1])
def test_gradient_finite_differences(capi, sbc):
    """dJ/dalpha and dJ/da . e (32 random directions e) against central
    differences of the GPU forward over a 6-step window with checkpoints
    every 4 steps (two segments, the second partial).  The differences are
    far above their rounding floor (about 1e-16 |J| / h)."""
    cfg = cfg_of("re180_64", 0.8, sbc, yplus=10.0)
    ch = capi.Channel(cfg)
    S0 = make_state(cfg, seed=21, amp=0.3)
    ch.set_state(S0)
    ch.init_stores()
    S0, RK0 = ch.get_state()
    rng = np.random.default_rng(5)
    act = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
    act[0, 0] = 0.0
    K, dt = 6, cfg.dt
    ch.set_control(0.8, act)
    ch.set_state(S0, RK0)
    J, dJa, g = ch.gradient(K, dt, ckpt_every=4, wq=0.5)
    assert J == pytest.approx(_J(ch, S0, RK0, 0.8, act, K, dt), rel=1e-14, abs=0)
    # central differences with eps = 1e-6 on the gain and on the actuation
    # (its std is 1e-3): a smaller step drowns the difference in the
    # rounding of the two forward runs
    h = 1e-6
    fd = (_J(ch, S0, RK0, 0.8 + h, act, K, dt) - _J(ch, S0, RK0, 0.8 - h, act, K, dt)) / (2 * h)
    print(f"J = {J:.12e}, dJ/dalpha = {dJa:.12e} (central difference {fd:.12e})")
    assert abs(fd) > 1e-6 * abs(J) and dJa == pytest.approx(fd, rel=1e-6), (dJa, fd)
    worst = 0.0
    for q in range(32):
        e = crand(rng, cfg.nz, cfg.nxh)
        e[0, 0] = 0.0
        he = 1e-6 / np.abs(e).max()
        fde = (_J(ch, S0, RK0, 0.8, act + he * e, K, dt) - _J(ch, S0, RK0, 0.8, act - he * e, K, dt)) / (2 * he)
        assert abs(fde) > 1e-6 * abs(J) and rdot(e, g) == pytest.approx(fde, rel=1e-6), (q, rdot(e, g), fde)
        worst = max(worst, abs(rdot(e, g) - fde) / abs(fde))
    print(f"dJ/da . e over 32 directions: worst relative difference {worst:.3e}; |dJ/da| = {np.abs(g).max():.3e}")
    ch.close()


def test_gradient_against_oracle_differences(capi, oracle_lib):
    """pb_gradient against central differences of the oracle's forward (the
    CPU restatement of the reference's step, with the same opposition gain
    and wall actuation) over a 4-step window, mixed-type scalar condition:
    dJ/dalpha and dJ/da . e for 4 random directions within 1e-5.  Steps:
    1e-5 on the gain and 1e-6 / max|e| on the actuation, where the oracle's
    own differences have converged to about 1e-7 and the device/oracle
    rounding difference (about 1e-13 |J|) stays below 1e-6 of them."""
    cfg = cfg_of("re180_64", 0.8, 1, yplus=10.0)
    ch = capi.Channel(cfg)
    ch.set_state(make_state(cfg, seed=21, amp=0.3))
    ch.init_stores()
    S0, RK0 = ch.get_state()
    rng = np.random.default_rng(5)
    act = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
    act[0, 0] = 0.0
    K, dt = 4, cfg.dt
    ch.set_control(0.8, act)
    ch.set_state(S0, RK0)
    J, dJa, g = ch.gradient(K, dt, ckpt_every=3, wq=0.5)
    ch.close()
    orc = oracle_lib.Oracle(cfg)

    def Jo(alpha, a):
        orc.set_control(alpha, a)
        S, RK = S0.copy(), RK0.copy()
        j = 0.0
        for _ in range(K):
            o = orc.step(S, RK, dt)[-1]
            j += o["tau_w"] + 0.5 * o["q_w"]
        return j / K

    assert J == pytest.approx(Jo(0.8, act), rel=1e-10)
    h = 1e-5
    fd = (Jo(0.8 + h, act) - Jo(0.8 - h, act)) / (2 * h)
    print(f"dJ/dalpha = {dJa:.12e}, oracle central difference {fd:.12e}")
    assert dJa == pytest.approx(fd, rel=1e-5)
    for q in range(4):
        e = crand(rng, cfg.nz, cfg.nxh)
        e[0, 0] = 0.0
        he = 1e-6 / np.abs(e).max()
        fde = (Jo(0.8, act + he * e) - Jo(0.8, act - he * e)) / (2 * he)
        print(f"direction {q}: dJ/da . e = {rdot(e, g):.12e}, oracle central difference {fde:.12e}")
        assert rdot(e, g) == pytest.approx(fde, rel=1e-5), q


def test_gradient_autograd_harness(capi):
    """The gradient wrapped as a torch.autograd.Function (test harness only):
    backward returns pb_gradient's values scaled by the incoming gradient."""
    import torch

    cfg = cfg_of("tiny", 0.5, 1, yplus=20.0)
    ch = capi.Channel(cfg)
    S0 = make_state(cfg, seed=2, amp=0.3)
    ch.set_state(S0)
    ch.init_stores()
    S0, RK0 = ch.get_state()
    K, dt = 4, cfg.dt

    class WallObjective(torch.autograd.Function):
        @staticmethod
        def forward(ctx, alpha, act):
            a = act.detach().cpu().numpy().view(np.complex128).reshape(cfg.nz, cfg.nxh)
            ch.set_control(float(alpha), a)
            ch.set_state(S0, RK0)
            J, dJa, g = ch.gradient(K, dt, ckpt_every=2)
            ctx.save_for_backward(torch.tensor(dJa, dtype=torch.float64),
                                  torch.from_numpy(g.view(np.float64).reshape(act.shape).copy()))
            return torch.tensor(J, dtype=torch.float64)

        @staticmethod
        def backward(ctx, gout):
            dJa, g = ctx.saved_tensors
            return gout * dJa, gout * g

    rng = np.random.default_rng(9)
    a0 = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
    a0[0, 0] = 0.0
    alpha = torch.tensor(0.5, dtype=torch.float64, requires_grad=True)
    act = torch.from_numpy(a0.view(np.float64).reshape(cfg.nz, 2 * cfg.nxh).copy()).requires_grad_(True)
    J = WallObjective.apply(alpha, act)
    (2.0 * J).backward()
    h = 1e-4
    Jp = _J(ch, S0, RK0, 0.5 + h, a0, K, dt)
    Jm = _J(ch, S0, RK0, 0.5 - h, a0, K, dt)
    assert alpha.grad.item() == pytest.approx(2.0 * (Jp - Jm) / (2 * h), rel=1e-6)
    assert act.grad.shape == act.shape and torch.isfinite(act.grad).all()
    ch.close()


def test_gradient_saved_state_paths_agree(capi, monkeypatch):
    """The reverse pass from saved z inverses and from saved states (the
    low-memory path) gives the same gradient bit for bit."""
    cfg = cfg_of("tiny", 0.7, 1, yplus=20.0)
    ch = capi.Channel(cfg)
    S0 = make_state(cfg, seed=8, amp=0.3)
    ch.set_state(S0)
    ch.init_stores()
    S0, RK0 = ch.get_state()
    rng = np.random.default_rng(4)
    act = 1e-3 * crand(rng, cfg.nz, cfg.nxh)
    ch.set_control(0.7, act)
    res = []
    for mode in ("", "state"):
        monkeypatch.setenv("PB_ADJ_SAVE", mode)
        ch.set_state(S0, RK0)
        res.append(ch.gradient(5, cfg.dt, ckpt_every=2))
    (Ja, da, ga), (Jb, db, gb) = res
    assert Ja == Jb and da == db and np.array_equal(ga, gb)
    ch.close()
