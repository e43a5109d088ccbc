"""Pins of the oracle's per-line physics by mathematics, not by its own formulas.

1. hvhg (hvhg.f90:141-276): h_g, h_v and h_theta are the wall-normal
   vorticity and the Laplacian of wall-normal velocity terms of the
   convective form H_i = -d_j(u_i u_j) (Kim, Moin & Moser 1987):
     h_g = d_z H_x - d_x H_z,
     h_v = -d_y (d_x H_x + d_z H_z) + (d_x^2 + d_z^2) H_y,
     h_theta = -d_j(u_j theta).
   The test builds H from separate uu, vv, ww (vv arbitrary: the
   P1 = uu - vv, P3 = ww - vv inputs must make it drop out) with polynomial
   y-profiles of degree <= 7, which the order-8 collocation differentiates
   exactly, and compares the oracle's hvhg with these definitions.
2. The wall-normal solves of one RK3-CN step (solve_g.f90, solve_t.f90,
   solve_v_f.f90, solve_u_w.f90, wave00.f90): three substeps of the oracle's
   wavespace on random products and stores against the textbook low-storage
   scheme of Spalart, Moser & Rogers (1991),
     (I - b_s L) q^s = (I + a_s L) q^(s-1) + gamma_s dt h^(s-1) + zeta_s dt h^(s-2),
   written as dense collocation systems: the phi/v pair as one coupled
   system (the phi equation on interior rows, phi = Laplacian(v) on interior
   rows, v = v_wall and v' = 0 at both walls) with no influence matrix and no
   mirror symmetry; u and w from continuity and the definition of g; the
   mean lines with the flow-rate (bulk) constraint as one extra unknown.
3. The whole substep in the linear limit: a small Orr-Sommerfeld eigenmode
   on laminar Poiseuille flow grows at the eigenvalue of a dense
   generalized eigenproblem (which itself reproduces Orszag's 1971 value
   c = 0.23752649 + 0.00373967 i at Re = 10^4, alpha = 1); for an oblique
   wave the Squire response g follows the same exponential.

Each pin is shown to fail under planted mistakes of the oracle
(oracle.mutant, the list is in oracle.cpp).
"""
import dataclasses

import numpy as np
import numpy.polynomial.polynomial as npoly
import pytest
import scipy.linalg as sl

from channel_passive_amd.config import ChannelConfig, physical_wavenumbers, preset
from tests.helpers import unpack_band

# Spalart, Moser & Rogers (1991), J. Comput. Phys. 96, 297: the low-storage
# RK3 / Crank-Nicolson coefficients, substeps 1..3
SMR_ALPHA = (29.0 / 96.0, -3.0 / 40.0, 1.0 / 6.0)
SMR_BETA = (37.0 / 160.0, 5.0 / 24.0, 1.0 / 6.0)
SMR_GAMMA = (8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0)
SMR_ZETA = (0.0, -17.0 / 60.0, -5.0 / 12.0)


def test_smr_coefficients_consistent():
    """alpha_s + beta_s = gamma_s + zeta_s (each substep advances the implicit
    and explicit parts by the same fraction of dt) and the fractions sum to 1."""
    for s in range(3):
        assert abs(SMR_ALPHA[s] + SMR_BETA[s] - SMR_GAMMA[s] - SMR_ZETA[s]) < 1e-15
    assert abs(sum(SMR_GAMMA) + sum(SMR_ZETA) - 1.0) < 1e-15


# ----------------------------------------------------------------- 1. hvhg
def _poly(rng, deg=7):
    return rng.standard_normal(deg + 1) + 1j * rng.standard_normal(deg + 1)


def hvhg_case(oracle_lib, cfg, seed, deg=7):
    orc = oracle_lib.Oracle(cfg)
    y = orc.y
    rng = np.random.default_rng(seed)
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    P = np.zeros((10, cfg.nz, cfg.nxh, cfg.ny), complex)
    want = {}
    ev = lambda c: npoly.polyval(y, c)
    d = lambda c: np.append(npoly.polyder(c), 0.0)  # same coefficient length
    for j in range(cfg.nz):
        if j == cfg.nz // 2:
            continue
        for k in range(cfg.nxh):
            if j == 0 and k == 0:
                continue
            kx, kz = kxp[k], kzp[j]
            uu, vv, ww, uv, vw, uw, ut, vt, wt = (_poly(rng, deg) for _ in range(9))
            P[0, j, k] = ev(uu - vv)
            P[1, j, k] = ev(uv)
            P[2, j, k] = ev(ww - vv)
            P[3, j, k] = ev(vw)
            P[4, j, k] = ev(uw)
            P[6, j, k] = ev(ut)
            P[7, j, k] = ev(vt)
            P[8, j, k] = ev(wt)
            # H_i = -(d_x u_i u + d_y u_i v + d_z u_i w), as polynomials in y
            Hx = -(1j * kx * uu + d(uv) + 1j * kz * uw)
            Hy = -(1j * kx * uv + d(vv) + 1j * kz * vw)
            Hz = -(1j * kx * uw + d(vw) + 1j * kz * ww)
            hg = 1j * kz * Hx - 1j * kx * Hz
            hv = -d(1j * kx * Hx + 1j * kz * Hz) - (kx * kx + kz * kz) * Hy
            ht = -(1j * kx * ut + d(vt) + 1j * kz * wt)
            want[(j, k)] = (ev(hg), ev(hv), ev(ht))
    return orc, P, want


def hvhg_error(orc, P, want):
    hg, hv, ht = orc.hvhg(P)
    worst = 0.0
    for (j, k), (wg, wv, wt) in want.items():
        for got, exp in ((hg[j, k], wg), (hv[j, k], wv), (ht[j, k], wt)):
            worst = max(worst, np.abs(got - exp).max() / np.abs(exp).max())
    return worst


# the collocation derivatives D1 LU(D0)^-1 and D2 LU(D0)^-1 of degree-7
# profiles on stretched knots carry roundoff of a few 1e-11 of the largest
# value of a line; a wrong term moves it by O(1)
HVHG_TOL = 1e-9


@pytest.mark.parametrize("name", ["tiny", "tiny_sin"])
def test_hvhg_from_definitions(oracle_lib, name):
    cfg, deg = preset("tiny"), 7
    if name == "tiny_sin":
        # the reference's sin-mapped knots with LABS_K = 6 (GSL order 7),
        # which reproduces polynomials of degree <= 6
        cfg, deg = dataclasses.replace(cfg, knot_map=0, stretch=0.9, labs_k=6, ny=41), 6
    orc, P, want = hvhg_case(oracle_lib, cfg, seed=3, deg=deg)
    err = hvhg_error(orc, P, want)
    print(f"{name}: hvhg worst relative error {err:.2e}")
    assert err <= HVHG_TOL


@pytest.mark.parametrize("m", [1, 2])
def test_hvhg_pin_fails_on_planted_mistake(oracle_lib, m):
    from oracle import oracle as O
    orc, P, want = hvhg_case(oracle_lib, preset("tiny"), seed=3)
    with O.mutant(m):
        err = hvhg_error(orc, P, want)
    assert err > 1e6 * HVHG_TOL, (m, err)


# ----------------------------------------------- 2. dense RK3-CN recurrence
def dense_ops(orc):
    D0, D1, D2 = (unpack_band(b, orc.bw) for b in (orc.d0, orc.d1, orc.d2))
    return D0, D1, D2, np.linalg.inv(D0)


def helmholtz_dense(D0, D2, k2, b, rhs):
    """c with rows 1..n-2 of ((1 + b k^2) D0 - b D2) c = rhs and q = D0 c = 0 at
    both walls."""
    n = D0.shape[0]
    M = (1.0 + b * k2) * D0 - b * D2
    M[0], M[-1] = D0[0], D0[-1]
    r = rhs.copy()
    r[0] = r[-1] = 0.0
    return np.linalg.solve(M, r)


def phi_v_dense(D0, D1, D2, k2, b, rhs, vlo, vhi):
    """The coupled phi/v problem: rows 1..n-2 of the phi equation and of
    phi = (D2 - k^2 D0) c_v, plus v = v_wall, v' = 0 at both walls."""
    n = D0.shape[0]
    A = np.zeros((2 * n, 2 * n))
    r = np.zeros(2 * n, complex)
    H = (1.0 + b * k2) * D0 - b * D2
    Q = D2 - k2 * D0
    q = 0
    for i in range(1, n - 1):
        A[q, :n] = H[i]
        r[q] = rhs[i]
        q += 1
    for i in range(1, n - 1):
        A[q, n:] = Q[i]
        A[q, :n] = -D0[i]
        q += 1
    for row, val in ((D0[0], vlo), (D0[-1], vhi), (D1[0], 0.0), (D1[-1], 0.0)):
        A[q, n:] = row
        r[q] = val
        q += 1
    x = np.linalg.solve(A, r)
    return x[:n], x[n:]


def mean_dense(D0, D2, w, b, rhs, bulk):
    """Mean line: rows 1..n-2 of (D0 - b D2) c - f = rhs with a uniform source
    f, c zero at both walls, and (if bulk is not None) 0.5 w.c = bulk."""
    n = D0.shape[0]
    M = D0 - b * D2
    if bulk is None:
        A = M.copy()
        A[0], A[-1] = D0[0], D0[-1]
        r = rhs.copy()
        r[0] = r[-1] = 0.0
        return np.linalg.solve(A, r), 0.0
    A = np.zeros((n + 1, n + 1))
    r = np.zeros(n + 1)
    A[:n, :n] = M
    A[1:n - 1, n] = -1.0
    r[:n] = rhs
    A[0, :n], A[n - 1, :n] = D0[0], D0[-1]
    A[0, n] = A[n - 1, n] = 0.0
    r[0] = r[n - 1] = 0.0
    A[n, :n] = 0.5 * w
    r[n] = bulk
    x = np.linalg.solve(A, r)
    return x[:n], x[n]


def dense_step_cfg(scalar_bc=0):
    # a viscosity large enough that the implicit terms move the result by
    # O(1e-2) of the fields; opposition control moves the walls
    return dataclasses.replace(preset("tiny"), nx=16, nz=8, ny=33, nu_value=0.02, dt=0.01, opp_alpha=0.6,
                               opp_yplus=30.0, bulk_u=0.9, bulk_t=1.3, scalar_bc=scalar_bc, name="dense")


def run_dense_check(oracle_lib, cfg, seed=17):
    """Three substeps of the oracle's wavespace against the dense recurrence;
    returns the worst relative error over every compared quantity."""
    orc = oracle_lib.Oracle(cfg)
    D0, D1, D2, D0i = dense_ops(orc)
    n, dt, nu, pr = cfg.ny, cfg.dt, cfg.nu, cfg.pr
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    rng = np.random.default_rng(seed)
    cplx = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)
    nh = cfg.nz // 2
    RK = cplx(3, cfg.nz, cfg.nxh, n)
    RK[:, nh] = 0.0
    RK[:, 0, 0] = RK[:, 0, 0].real  # the mean-line stores are real
    Ps = []
    for s in range(3):
        P = cplx(10, cfg.nz, cfg.nxh, n)
        P[:, nh] = 0.0
        P[:, 0, 0] = P[:, 0, 0].real
        Ps.append(P)
    vd = cplx(2, cfg.nz, cfg.nxh)
    vd[:, nh] = 0.0
    # mixed-type scalar condition: the source (nu/Pr) u / U_b with u of each
    # substep's input state (real on the mean line)
    mixed = cfg.scalar_bc == 1
    ks = nu / pr / cfg.bulk_u_value()
    us = [cplx(cfg.nz, cfg.nxh, n) for _ in range(3)]
    for u in us:
        u[nh] = 0.0
        u[0, 0] = u[0, 0].real
    hs = [orc.hvhg(P) for P in Ps]  # pinned by test_hvhg_from_definitions
    dy = D1 @ D0i
    d2 = D2 @ D0i
    # q^(s-1) per line: (g, phi, theta), on the mean line (U, W, Theta); at
    # s = 1 the stores hold the state's values
    q = {(j, k): [RK[0, j, k].copy(), RK[1, j, k].copy(), RK[2, j, k].copy()] for j in range(cfg.nz)
         for k in range(cfg.nxh) if j != nh}
    worst = 0.0

    def cmp(got, exp, what):
        nonlocal worst
        e = np.abs(got - exp).max() / max(np.abs(exp).max(), 1e-300)
        worst = max(worst, e)
        return e

    RKo = RK.copy()
    for s in (1, 2, 3):
        a_s, b_s, c_s = SMR_ALPHA[s - 1] * dt, SMR_BETA[s - 1] * dt, SMR_GAMMA[s - 1] * dt
        z_s = SMR_ZETA[s - 1] * dt
        U = Ps[s - 1].copy()
        mean = orc.wavespace(s, dt, U, RKo, vd, usrc=us[s - 1] if mixed else None)
        for (j, k), (g0, p0, t0) in q.items():
            kx, kz = kxp[k], kzp[j]
            k2 = kx * kx + kz * kz
            hp = [None, None, None] if s == 1 else [hs[s - 2][0][j, k], hs[s - 2][1][j, k], hs[s - 2][2][j, k]]
            if j == 0 and k == 0:
                # mean lines: h_U = -d(uv)/dy, h_W = -d(vw)/dy, h_T = -d(v theta)/dy
                out = []
                for f, (slot, fac, bulk, src) in enumerate(((1, 1.0, cfg.bulk_u_value(), 0),
                                                             (3, 1.0, None, 2),
                                                             (7, 1.0 / pr, None if mixed else cfg.bulk_t, 6))):
                    qq = (g0, p0, t0)[f].real
                    h = -(dy @ Ps[s - 1][slot, 0, 0].real)
                    hprev = 0.0 if s == 1 else -(dy @ Ps[s - 2][slot, 0, 0].real)
                    if f == 2 and mixed:  # heat-flux source on the mean scalar
                        h = h + ks * us[s - 1][0, 0].real
                        hprev = hprev if s == 1 else hprev + ks * us[s - 2][0, 0].real
                    rhs = qq + a_s * nu * fac * (d2 @ qq) + c_s * h + z_s * hprev
                    c, force = mean_dense(D0, D2, orc.integ, b_s * nu * fac, rhs, bulk)
                    new = D0 @ c
                    cmp(U[src, 0, 0], new, f"mean {f} s{s}")
                    if f == 0:
                        cmp(np.array([mean[0]]), np.array([force]), "phi_u")
                    if f == 2:
                        cmp(np.array([mean[1]]), np.array([force]), "phi_t")
                    if s < 3:
                        a1, z1 = SMR_ALPHA[s] * dt * nu * fac, SMR_ZETA[s] * dt
                        st = new + a1 * (D2 @ c) + z1 * h
                    else:
                        st = new
                    cmp(RKo[f, 0, 0], st, f"mean store {f} s{s}")
                    out.append(new + 0j)
                q[(j, k)] = out
                continue
            hg, hv, ht = hs[s - 1][0][j, k], hs[s - 1][1][j, k], hs[s - 1][2][j, k]
            if mixed:
                ht = ht + ks * us[s - 1][j, k]
                if s > 1:
                    hp[2] = hp[2] + ks * us[s - 2][j, k]
            new_vals, stores = [], []
            for f, (q0, h, fac) in enumerate(((g0, hg, 1.0), (t0, ht, 1.0 / pr))):
                rhs = q0 + a_s * nu * fac * (d2 @ q0 - k2 * q0) + c_s * h
                if s > 1:
                    rhs = rhs + z_s * hp[2 * f]
                c = helmholtz_dense(D0, D2, k2, b_s * nu * fac, rhs)
                new = D0 @ c
                if s < 3:
                    a1, z1 = SMR_ALPHA[s] * dt * nu * fac, SMR_ZETA[s] * dt
                    st = new + a1 * (D2 @ c - k2 * new) + z1 * h
                else:
                    st = new
                new_vals.append(new)
                stores.append(st)
            g_new, t_new = new_vals
            rhs = p0 + a_s * nu * (d2 @ p0 - k2 * p0) + c_s * hv
            if s > 1:
                rhs = rhs + z_s * hp[1]
            vlo, vhi = -cfg.opp_alpha * vd[0, j, k], -cfg.opp_alpha * vd[1, j, k]
            cphi, cv = phi_v_dense(D0, D1, D2, k2, b_s * nu, rhs, vlo, vhi)
            phi_new = D0 @ cphi
            v_new, f_new = D0 @ cv, D1 @ cv
            if s < 3:
                a1, z1 = SMR_ALPHA[s] * dt * nu, SMR_ZETA[s] * dt
                phi_st = phi_new + a1 * (D2 @ cphi - k2 * phi_new) + z1 * hv
            else:
                phi_st = phi_new
            # u, w from continuity (i kx u + v' + i kz w = 0) and the
            # definition of the wall-normal vorticity (i kz u - i kx w = g)
            M2 = np.array([[1j * kx, 1j * kz], [1j * kz, -1j * kx]])
            uw = np.linalg.solve(M2, np.vstack([-f_new, g_new]))
            cmp(U[1, j, k], v_new, "v")
            cmp(U[0, j, k], uw[0], "u")
            cmp(U[2, j, k], uw[1], "w")
            cmp(U[6, j, k], t_new, "theta")
            cmp(RKo[0, j, k], stores[0], "store g")
            cmp(RKo[1, j, k], phi_st, "store phi")
            cmp(RKo[2, j, k], stores[1], "store theta")
            q[(j, k)] = [g_new, phi_new, t_new]
    return worst


DENSE_TOL = 1e-10


@pytest.mark.parametrize("scalar_bc", [0, 1])
def test_wall_normal_solves_dense_recurrence(oracle_lib, scalar_bc):
    """scalar_bc 1: the mixed-type scalar condition, h_theta + (nu/Pr) u/U_b
    on every line and no held bulk on the mean scalar."""
    err = run_dense_check(oracle_lib, dense_step_cfg(scalar_bc))
    print(f"dense RK3-CN recurrence (scalar_bc {scalar_bc}): worst relative error {err:.2e}")
    assert err <= DENSE_TOL


@pytest.mark.parametrize("m,scalar_bc", [(3, 0), (4, 0), (5, 0), (6, 0), (7, 0), (8, 0), (9, 1)])
def test_dense_recurrence_fails_on_planted_mistake(oracle_lib, m, scalar_bc):
    from oracle import oracle as O
    with O.mutant(m):
        err = run_dense_check(oracle_lib, dense_step_cfg(scalar_bc))
    assert err > 1e3 * DENSE_TOL, (m, err)


# ----------------------------------------------- 3. Orr-Sommerfeld / Squire
def os_operators(orc, kx, kz, nu):
    """Least stable Orr-Sommerfeld eigenvalue of U = 1 - y^2 in the
    collocation space, lambda phi = -i kx U phi + i kx U'' v + nu Lap(phi)
    on interior rows, phi = Lap(v) on interior rows, v = v' = 0 at the walls;
    and the forced Squire response lambda g = -i kx U g - i kz U' v + nu Lap(g),
    g = 0 at the walls.  Returns (lambda, v, v', g) as collocation values."""
    n, y = orc.cfg.ny, orc.y
    D0, D1, D2, _ = dense_ops(orc)
    k2 = kx * kx + kz * kz
    U, Up, Upp = 1 - y * y, -2 * y, -2.0 * np.ones(n)
    Lap = D2 - k2 * D0
    A = np.zeros((2 * n, 2 * n), complex)
    B = np.zeros((2 * n, 2 * n))
    r = 0
    for i in range(1, n - 1):
        A[r, n:] = -1j * kx * U[i] * D0[i] + nu * Lap[i]
        A[r, :n] = 1j * kx * Upp[i] * D0[i]
        B[r, n:] = D0[i]
        r += 1
    for i in range(1, n - 1):
        A[r, :n] = Lap[i]
        A[r, n:] = -D0[i]
        r += 1
    for row in (D0[0], D0[-1], D1[0], D1[-1]):
        A[r, :n] = row
        r += 1
    w, V = sl.eig(A, B)
    ok = np.isfinite(w) & (np.abs(w) < 1e3)
    i = np.flatnonzero(ok)[np.argmax(w[ok].real)]
    lam, cv = w[i], V[:n, i]
    M = lam * D0 + 1j * kx * U[:, None] * D0 - nu * Lap
    rhs = -1j * kz * Up * (D0 @ cv)
    M[0], M[-1] = D0[0], D0[-1]
    rhs[0] = rhs[-1] = 0.0
    cg = np.linalg.solve(M, rhs)
    return lam, D0 @ cv, D1 @ cv, D0 @ cg


def os_cfg(kzi):
    # Re = U_c h / nu = 10^4 on laminar U = 1 - y^2 (bulk 2/3); Lx = 2 pi
    # (alpha = 1 at kx index 1), Lz = pi (beta = 2 at kz index 1)
    return ChannelConfig(16, 97, 4 if kzi == 0 else 8, re_tau=180.0, lmfx=1.0, lmfz=0.5, dt=0.01, nu_value=1e-4,
                         bulk_u=2.0 / 3.0, name="os")


def os_run(oracle_lib, kzi, T=3.0, eps=1e-7, mut=0):
    from oracle import oracle as O
    cfg = os_cfg(kzi)
    orc = oracle_lib.Oracle(cfg)
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    kx, kz = kxp[1], kzp[kzi]
    lam, v, f, g = os_operators(orc, kx, kz, cfg.nu)
    y = orc.y
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), complex)
    S[0, 0, 0] = 1 - y * y
    S[3, 0, 0] = 1.5 * (1 - y * y)
    k2 = kx * kx + kz * kz
    u, w = (1j * kx * f - 1j * kz * g) / k2, (1j * kz * f + 1j * kx * g) / k2
    sc = eps / max(np.abs(a).max() for a in (u, v, w))
    S[0, kzi, 1], S[1, kzi, 1], S[2, kzi, 1] = u * sc, v * sc, w * sc
    RK = orc.init_stores(S)
    N = int(round(T / cfg.dt))
    with O.mutant(mut):
        for _ in range(N):
            orc.step(S, RK, cfg.dt)
    grow = np.exp(lam * N * cfg.dt)
    vT = S[1, kzi, 1]
    gT = 1j * kz * S[0, kzi, 1] - 1j * kx * S[2, kzi, 1]
    v_res = np.abs(vT - grow * v * sc).max() / np.abs(vT).max()
    g_res = np.abs(gT - grow * g * sc).max() / np.abs(gT).max() if kz != 0 else 0.0
    lam_v = np.log(np.vdot(v, vT) / np.vdot(v, v * sc)) / (N * cfg.dt)
    return lam, lam_v, v_res, g_res


def test_orr_sommerfeld_dense_matches_orszag(oracle_lib):
    """The dense collocation eigenproblem reproduces Orszag (1971),
    J. Fluid Mech. 50, 689: c = 0.23752649 + 0.00373967 i at Re = 10^4,
    alpha = 1 (lambda = -i alpha c)."""
    orc = oracle_lib.Oracle(os_cfg(0))
    lam, _, _, _ = os_operators(orc, 1.0, 0.0, 1e-4)
    c = 1j * lam
    assert abs(c.real - 0.23752649) <= 5e-8 and abs(c.imag - 0.00373967) <= 5e-8, c


def test_orr_sommerfeld_growth_2d(oracle_lib):
    """Three time units of the full oracle step from a 2-D eigenmode of
    amplitude 1e-7: v grows and travels at the eigenvalue (the RK3-CN error
    at dt = 0.01 is ~1e-9 here, the nonlinear terms ~1e-7 relative)."""
    lam, lam_v, v_res, _ = os_run(oracle_lib, 0)
    print(f"2-D OS: lambda {lam:.10f}, from the steps {lam_v:.10f}, v residual {v_res:.1e}")
    assert abs(lam_v - lam) <= 1e-7 and v_res <= 1e-6


def test_orr_sommerfeld_squire_oblique(oracle_lib):
    """Oblique wave alpha = 1, beta = 2: v at the Orr-Sommerfeld eigenvalue and
    the wall-normal vorticity g (driven by -i beta U' v through h_g) at the same
    exponential; the bar covers RK3-CN's O(dt^2) error at |lambda dt| ~ 0.01."""
    lam, lam_v, v_res, g_res = os_run(oracle_lib, 1)
    print(f"oblique OS: lambda {lam:.8f}, from the steps {lam_v:.8f}, v residual {v_res:.1e}, g residual {g_res:.1e}")
    assert abs(lam_v.real - lam.real) <= 1e-5 and v_res <= 1e-3 and g_res <= 1e-3


@pytest.mark.parametrize("m,kzi", [(1, 1), (2, 0), (7, 0)])
def test_orr_sommerfeld_fails_on_planted_mistake(oracle_lib, m, kzi):
    """The growth tests above reject a wrong h_g term (oblique wave), a
    dropped h_v term and a wrong sign of b k^2 in the Helmholtz operator
    (the last moves v by ~1e-4 at nu = 1e-4: 100 x the 2-D bar)."""
    lam, lam_v, v_res, g_res = os_run(oracle_lib, kzi, mut=m)
    if kzi == 0:
        assert abs(lam_v - lam) > 1e-7 or v_res > 1e-6, (m, lam_v, v_res)
    else:
        assert abs(lam_v.real - lam.real) > 1e-5 or v_res > 1e-3 or g_res > 1e-3, (m, v_res, g_res)
