"""Seeded, divergence-free initial states in the reference spectral layout.

The reference starts from a fixed three-mode field
(SOURCE_PASSOPB/initial_step.f90:63-121); the benchmark workloads instead
start from a laminar (or turbulent-mean) profile plus a random,
divergence-free perturbation built from a streamfunction pair:

* v-hat = A a_v (1-y^2)^2          (v = v' = 0 at the walls)
* g-hat = A k a_g (1-y^2)          (wall-normal vorticity, 0 at the walls)
* u, w from continuity exactly as solve_u_w (solve_u_w.f90:27-38):
  u = (i kx f - i kz g)/k^2,  w = (i kz f + i kx g)/k^2,  f = dv/dy
* theta' = A a_t (1-y^2)

for index wavenumbers 0 <= kx <= K, |kz| <= K, (kx,kz) != (0,0), K = 8,
with a_* complex Gaussians from a splitmix64 + Box-Muller stream (so every
consumer that is handed the same seed gets the same bits), Hermitian
symmetry on the kx = 0 plane and the kz Nyquist slot left at zero.

Arrays are numpy complex128 of shape (nslot, nz, nx/2, ny) in C order, which
is the reference's Fortran ``U(ny, nx/2, nz, nslot)``.
"""
from __future__ import annotations

import math

import numpy as np

from .config import PB_SCALAR_MIXED, ChannelConfig, physical_wavenumbers

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 (Steele, Lea, Flood 2014) with Box-Muller normals."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        return ((self.next_u64() >> 11) + 0.5) * (1.0 / 9007199254740992.0)

    def normal_pair(self):
        u1, u2 = self.uniform(), self.uniform()
        r = math.sqrt(-2.0 * math.log(u1))
        return r * math.cos(2.0 * math.pi * u2), r * math.sin(2.0 * math.pi * u2)

    def cnormal(self) -> complex:
        a, b = self.normal_pair()
        return complex(a, b)


def breakpoints(cfg: ChannelConfig) -> np.ndarray:
    """Breakpoints of the B-spline basis (labs_fill_full.cpp:80-101)."""
    k = cfg.labs_k + 1
    nbreak = cfg.ny - k + 2
    i = np.arange(nbreak, dtype=np.float64)
    ly = 2.0 * cfg.lmfy
    if cfg.knot_map == 1:
        return 0.5 * ly * np.tanh(cfg.stretch * (-1.0 + (2.0 * i) / (nbreak - 1))) / np.tanh(cfg.stretch)
    if abs(cfg.stretch) <= 1e-16:
        return -1.0 + 2.0 / (nbreak - 1) * i
    pi = 3.141592653589793238462643
    return np.sin(cfg.stretch * pi / 2.0 * (-1.0 + (2.0 * i) / (nbreak - 1))) / np.sin(cfg.stretch * pi / 2.0)


def collocation_points(cfg: ChannelConfig) -> np.ndarray:
    """Greville abscissae of the order-(labs_k+1) basis: running mean of
    k-1 interior knots in extended precision (as gsl_stats_mean does)."""
    k = cfg.labs_k + 1
    brk = breakpoints(cfg)
    knots = np.concatenate([np.full(k - 1, brk[0]), brk, np.full(k - 1, brk[-1])])
    y = np.empty(cfg.ny)
    for i in range(cfg.ny):
        mean = np.longdouble(0.0)
        for m in range(k - 1):
            mean += (np.longdouble(knots[i + 1 + m]) - mean) / np.longdouble(m + 1)
        y[i] = float(mean)
    return y


def mean_profile(cfg: ChannelConfig, y: np.ndarray, turbulent: bool) -> np.ndarray:
    ub = cfg.bulk_u_value()
    if not turbulent:
        return 1.5 * ub * (1.0 - y * y)
    yp = (1.0 - np.abs(y)) * cfg.re_tau
    kappa = 0.41
    u = np.log1p(kappa * yp) / kappa + 7.8 * (1.0 - np.exp(-yp / 11.0) - (yp / 11.0) * np.exp(-yp / 3.0))
    bulk = np.trapezoid(u, y) / (y[-1] - y[0])
    return u * (ub / bulk)


def scalar_mean_profile(cfg: ChannelConfig, y: np.ndarray) -> np.ndarray:
    """Laminar scalar profile of the configuration's forcing: 1.5 Theta_b
    (1 - y^2) under the held-bulk uniform source, (5 - 6 y^2 + y^4) / 8 under
    uniform wall heat flux with the mixed-type condition (the steady solution
    of (nu/Pr) Theta'' = -(nu/Pr) U/U_b on U = 1.5 U_b (1 - y^2))."""
    if cfg.scalar_bc == PB_SCALAR_MIXED:
        y2 = y * y
        return (5.0 - 6.0 * y2 + y2 * y2) / 8.0
    return 1.5 * cfg.bulk_t * (1.0 - y * y)


def make_state(cfg: ChannelConfig, seed: int = 0xC0FFEE, amp: float = 0.1, kmax: int = 8,
               turbulent_mean: bool | None = None) -> np.ndarray:
    """State S(4, nz, nx/2, ny): u, v, w, theta collocation values."""
    if turbulent_mean is None:
        turbulent_mean = cfg.re_tau >= 500.0
    y = collocation_points(cfg)
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), dtype=np.complex128)
    U = mean_profile(cfg, y, turbulent_mean)
    S[0, 0, 0, :] = U
    S[3, 0, 0, :] = scalar_mean_profile(cfg, y)
    _, _, kxp, kzp = physical_wavenumbers(cfg)
    nh = cfg.nz // 2
    rng = SplitMix64(seed)
    modes = []
    for ix in range(0, min(kmax, cfg.nxh - 1) + 1):
        for iz in range(-min(kmax, nh - 1), min(kmax, nh - 1) + 1):
            if ix == 0 and iz <= 0:
                continue
            modes.append((ix, iz))
    if not modes:
        return S
    A = amp * float(np.max(U)) / math.sqrt(len(modes))
    w2 = 1.0 - y * y
    for ix, iz in modes:
        j = iz if iz >= 0 else cfg.nz + iz
        kx, kz = kxp[ix], kzp[j]
        k2 = kx * kx + kz * kz
        k = math.sqrt(k2)
        av, ag, at = rng.cnormal(), rng.cnormal(), rng.cnormal()
        v = A * av * w2 * w2
        f = A * av * (-4.0 * y * w2)
        g = A * k * ag * w2
        u = (1j * kx * f - 1j * kz * g) / k2
        w = (1j * kz * f + 1j * kx * g) / k2
        th = 0.1 * A * at * w2
        for slot, val in enumerate((u, v, w, th)):
            S[slot, j, ix, :] = val
            if ix == 0:
                S[slot, (cfg.nz - iz) % cfg.nz, 0, :] = np.conj(val)
    return S


def reference_state(cfg: ChannelConfig, y: np.ndarray) -> np.ndarray:
    """The reference's cold-start field (set_initial_value_in_y_pencil,
    initial_step.f90:63-121) on the collocation points y: the laminar
    profile 1.5 (1 - y^2) - u_gal on the mean mode, a v mode at kx index 1
    and a (u, v, w) mode pair at kz = +-1 on kx = 0 (conjugates of each
    other); theta is a copy of u on every mode.  State S(4, nz, nx/2, ny)."""
    S = np.zeros((4, cfg.nz, cfg.nxh, cfg.ny), dtype=np.complex128)
    y = np.asarray(y, dtype=np.float64)
    a = 1.0 - y ** 2
    b = ((1.0 - y) * (1.0 + y)) ** 2
    S[0, 0, 0] = a * 1.5 - cfg.u_gal
    if cfg.nxh > 1:
        S[1, 0, 1] = b * 2e-2 + 1j * (b * 1e-2)
    if cfg.nz > 1:
        S[0, 1, 0] = a * 2e-2 + 1j * (a * 1e-2)
        S[1, 1, 0] = -a * 1e-2 + 1j * (a * 1e-2)
        S[2, 1, 0] = -a * 1e-2 + 1j * (-a * 2e-2)
        S[0, cfg.nz - 1, 0] = a * 2e-2 + 1j * (-a * 1e-2)
        S[1, cfg.nz - 1, 0] = -a * 1e-2 + 1j * (-a * 1e-2)
        S[2, cfg.nz - 1, 0] = -a * 1e-2 + 1j * (a * 2e-2)
    S[3] = S[0]
    return S
