"""Discrete adjoint of the wall-normal stage (channel_passive_amd/csrc/solver_adj.h),
built for the CPU, against the forward (solver.h) by dot-product tests.

For a fixed dt the stage W maps (products, stores, detection-plane v, wall
actuation, mixed-type source u) affinely to (u, v, w, theta, stores, tau_w,
q_w); M x = W(x) - W(0) is exactly linear, so <M x, y> = <x, M^T y> holds to
rounding for random x and y under the real inner product sum Re(conj(a) b)
over every stored complex number.  d/d alpha is checked the same way: W is
affine in alpha at fixed inputs.  A planted mistake in the adjoint (a
transposed operand, a dropped term) breaks these identities at O(1).
"""
import ctypes
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset

dp = ctypes.POINTER(ctypes.c_double)
PROD = (0, 1, 2, 3, 4, 6, 7, 8)
OUT = (0, 1, 2, 6)


def rdot(a, b):
    return float(np.sum(a.real * b.real + a.imag * b.imag))


def crand(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def forward(emu_lib, cfg, s, U, RK, vd, usrc, act):
    c = cfg.to_c()
    U = np.ascontiguousarray(U.copy())
    RK = np.ascontiguousarray(RK.copy())
    m = np.zeros(4)
    ptr = lambda a: None if a is None else a.ctypes.data_as(dp)
    assert emu_lib.emu_wavespace_act(ctypes.byref(c), s, ctypes.c_double(cfg.dt), ptr(U), ptr(RK), ptr(vd),
                                     ptr(usrc), ptr(act), ptr(m)) == 0
    return U, RK, m


def adjoint(emu_lib, cfg, s, AU, ARK, vd, tau_bar, q_bar):
    c = cfg.to_c()
    plane = cfg.nz * (cfg.nx // 2)
    AU = np.ascontiguousarray(AU.copy())
    ARK = np.ascontiguousarray(ARK.copy())
    avd = np.zeros(2 * plane, complex)
    aact = np.zeros(plane, complex)
    aal = np.zeros(plane)
    assert emu_lib.emu_adj_wavespace(ctypes.byref(c), s, ctypes.c_double(cfg.dt), AU.ctypes.data_as(dp),
                                     ARK.ctypes.data_as(dp), None if vd is None else vd.ctypes.data_as(dp),
                                     avd.ctypes.data_as(dp), aact.ctypes.data_as(dp), aal.ctypes.data_as(dp),
                                     ctypes.c_double(tau_bar), ctypes.c_double(q_bar)) == 0
    return AU, ARK, avd, aact, aal


CASES = [("tiny", 0.0, 0), ("tiny", 0.8, 0), ("tiny", 0.6, 1), ("re180_64", 0.5, 0)]


@pytest.mark.parametrize("name,alpha,sbc", CASES)
@pytest.mark.parametrize("s", [1, 2, 3])
def test_wavespace_adjoint_dot(emu_lib, name, alpha, sbc, s):
    cfg = dataclasses.replace(preset(name), opp_alpha=alpha, opp_yplus=15.0, scalar_bc=sbc)
    nxh, nz, ny = cfg.nx // 2, cfg.nz, cfg.ny
    plane = nz * nxh
    rng = np.random.default_rng(100 * s + int(10 * alpha) + sbc)
    U = crand(rng, 9, ny, nz, nxh)
    U[5] = 0.0
    RK = crand(rng, 3, ny, nz, nxh)
    vd = crand(rng, 2 * plane)
    usrc = crand(rng, ny, nz, nxh)
    act = 1e-1 * crand(rng, plane)
    Ux, Rx, mx = forward(emu_lib, cfg, s, U, RK, vd, usrc, act)
    z = lambda a: np.zeros_like(a)
    U0, R0, m0 = forward(emu_lib, cfg, s, z(U), z(RK), z(vd), z(usrc), z(act))
    AU = crand(rng, 9, ny, nz, nxh)  # slots other than 0,1,2,6 hold junk the adjoint must ignore
    ARK = crand(rng, 3, ny, nz, nxh)
    tb, qb = rng.standard_normal(2)
    lhs = sum(rdot(Ux[f] - U0[f], AU[f]) for f in OUT) + rdot(Rx - R0, ARK)
    if s == 3:
        lhs += (mx[2] - m0[2]) * tb + (mx[3] - m0[3]) * qb
    PU, PR, avd, aact, aal = adjoint(emu_lib, cfg, s, AU, ARK, vd, tb if s == 3 else 0.0, qb if s == 3 else 0.0)
    rhs = sum(rdot(U[f], PU[f]) for f in PROD) + rdot(RK, PR) + rdot(act, aact)
    if alpha:
        rhs += rdot(vd, avd)
    if sbc == 1:
        rhs += rdot(usrc, PU[5])
    scale = np.sqrt(sum(rdot(Ux[f] - U0[f], Ux[f] - U0[f]) for f in OUT) * sum(rdot(AU[f], AU[f]) for f in OUT))
    assert abs(lhs - rhs) <= 1e-12 * scale, (lhs, rhs, abs(lhs - rhs) / scale)
    # d/d alpha at fixed inputs (vd given even at alpha = 0)
    ca = dataclasses.replace(cfg, opp_alpha=alpha + 0.5)
    cb = dataclasses.replace(cfg, opp_alpha=alpha - 0.5)
    Ua, Ra, ma = forward(emu_lib, ca, s, U, RK, vd, usrc, act)
    Ub, Rb, mb = forward(emu_lib, cb, s, U, RK, vd, usrc, act)
    dl = sum(rdot(Ua[f] - Ub[f], AU[f]) for f in OUT) + rdot(Ra - Rb, ARK)
    if s == 3:
        dl += (ma[2] - mb[2]) * tb + (ma[3] - mb[3]) * qb
    # the per-system terms cancel heavily: bound the rounding by their magnitudes
    assert abs(dl - aal.sum()) <= 1e-12 * (np.abs(aal).sum() + abs(dl)), (dl, aal.sum())


def test_dot_test_detects_a_wrong_transpose(emu_lib):
    """The identity is sharp: the transpose of another substep (other RK
    coefficients, a plausible slip) misses it at O(1)."""
    cfg = dataclasses.replace(preset("tiny"), opp_alpha=0.8, opp_yplus=15.0)
    nxh, nz, ny = cfg.nx // 2, cfg.nz, cfg.ny
    plane = nz * nxh
    rng = np.random.default_rng(1)
    U = crand(rng, 9, ny, nz, nxh)
    U[5] = 0.0
    RK = crand(rng, 3, ny, nz, nxh)
    vd = crand(rng, 2 * plane)
    Ux, Rx, _ = forward(emu_lib, cfg, 2, U, RK, vd, None, None)
    z = lambda a: np.zeros_like(a)
    U0, R0, _ = forward(emu_lib, cfg, 2, z(U), z(RK), z(vd), None, None)
    AU = crand(rng, 9, ny, nz, nxh)
    ARK = crand(rng, 3, ny, nz, nxh)
    lhs = sum(rdot(Ux[f] - U0[f], AU[f]) for f in OUT) + rdot(Rx - R0, ARK)
    for s_wrong in (1, 3):
        PU, PR, avd, _, _ = adjoint(emu_lib, cfg, s_wrong, AU, ARK, vd, 0.0, 0.0)
        rhs = sum(rdot(U[f], PU[f]) for f in PROD) + rdot(RK, PR) + rdot(vd, avd)
        assert abs(lhs - rhs) > 1e-3 * abs(lhs), (s_wrong, lhs, rhs)
