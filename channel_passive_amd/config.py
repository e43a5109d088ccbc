"""Run configurations for the channel time step and the ctypes mirror of
``pb_config`` (include/pb_capi.h).

The physical set-up follows the reference's ``input`` records
(SOURCE_PASSOPB/read_input_file.f90:39-116): box sizes are given as
``LMFx, LMFy, LMFz`` with Lx = 2*pi*LMFx, Ly = 2*LMFy, Lz = 2*pi*LMFz, the
flow rate and the bulk scalar are held at ``bulk_u`` / ``bulk_t``
(wave00.f90:202, 612).

Units are bulk units, as the reference's own cold start uses them
(initial_step.f90:83: laminar 1.5 (1 - y^2), i.e. U_b = 1, half height
h = 1): U_b = 1 and nu = 1 / Re_b with Re_b = U_b h / nu taken from Dean's
correlation Re_tau = 0.09 (2 Re_b)^0.88 for the named Re_tau.  At Re_tau =
180 this gives nu = 1/2819 and dt = 0.005 is dt+ = 0.057; the friction-unit
reading dt+ = 0.9 puts pi * dt * velmax near 9 on the 64 x 65 x 64 box,
far past RK3's imaginary-axis limit sqrt(3).

Each preset's fixed dt is chosen from the reference's CFL rule
dt <= CFL / (pi * velmax) (read_input_file.f90:107, dft_x_realspace.f90:492)
on its seeded initial state, with pi * dt * velmax between 0.45 and 0.65.
"""
from __future__ import annotations

import ctypes
import dataclasses
import math

PB_KNOTS_REFERENCE = 0
PB_KNOTS_TANH = 1
# scalar forcing: the reference's held-bulk uniform source, or uniform wall
# heat flux with the mixed-type condition (source (nu/Pr) u / U_b)
PB_SCALAR_SOURCE = 0
PB_SCALAR_MIXED = 1


class PbConfig(ctypes.Structure):
    """Byte-for-byte mirror of ``struct pb_config``."""

    _fields_ = [
        ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
        ("labs_k", ctypes.c_int32), ("knot_map", ctypes.c_int32),
        ("stretch", ctypes.c_double),
        ("lmfx", ctypes.c_double), ("lmfy", ctypes.c_double), ("lmfz", ctypes.c_double),
        ("nu", ctypes.c_double), ("pr", ctypes.c_double),
        ("bulk_u", ctypes.c_double), ("bulk_w", ctypes.c_double), ("bulk_t", ctypes.c_double),
        ("u_gal", ctypes.c_double),
        ("cfl", ctypes.c_double), ("dt_factor", ctypes.c_double), ("dy_factor", ctypes.c_double),
        ("opp_alpha", ctypes.c_double), ("opp_ydist", ctypes.c_double),
        ("rk_step", ctypes.c_int32), ("scalar_bc", ctypes.c_int32),
    ]


def dean_bulk_reynolds(re_tau: float) -> float:
    """Re_b = U_b h / nu from Dean's correlation Re_tau = 0.09 Re_2h^0.88,
    Re_2h = 2 h U_b / nu."""
    return 0.5 * (re_tau / 0.09) ** (1.0 / 0.88)


def dean_bulk_velocity(re_tau: float) -> float:
    """U_b / u_tau = Re_b / Re_tau (the bulk velocity in friction units)."""
    return dean_bulk_reynolds(re_tau) / re_tau


@dataclasses.dataclass
class ChannelConfig:
    nx: int
    ny: int
    nz: int
    re_tau: float = 180.0
    lmfx: float = 1.0
    lmfy: float = 1.0
    lmfz: float = 0.5
    labs_k: int = 7               # LABS_K; GSL order labs_k + 1
    knot_map: int = PB_KNOTS_TANH
    stretch: float = 2.0
    pr: float = 0.71
    bulk_u: float | None = None   # default 1 (bulk units)
    bulk_w: float = 0.0
    bulk_t: float = 1.0
    u_gal: float = 0.0
    cfl: float = 1.0
    dt_factor: float = 1.0        # fixed dt
    dy_factor: float = 1.0
    opp_alpha: float = 0.0
    opp_yplus: float = 10.0
    dt: float = 0.005
    rk_step: int = 3
    scalar_bc: int = PB_SCALAR_SOURCE
    name: str = ""

    nu_value: float | None = None  # default 1 / Re_b (Dean's correlation)

    @property
    def nu(self) -> float:
        return self.nu_value if self.nu_value is not None else 1.0 / dean_bulk_reynolds(self.re_tau)

    @property
    def nxh(self) -> int:
        return self.nx // 2

    @property
    def n_spec(self) -> int:
        return self.ny * self.nxh * self.nz

    def to_c(self) -> PbConfig:
        c = PbConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        c.labs_k, c.knot_map, c.stretch = self.labs_k, self.knot_map, self.stretch
        c.lmfx, c.lmfy, c.lmfz = self.lmfx, self.lmfy, self.lmfz
        c.nu, c.pr = self.nu, self.pr
        c.bulk_u = self.bulk_u_value()
        c.bulk_w, c.bulk_t, c.u_gal = self.bulk_w, self.bulk_t, self.u_gal
        c.cfl, c.dt_factor, c.dy_factor = self.cfl, self.dt_factor, self.dy_factor
        c.opp_alpha = self.opp_alpha
        c.opp_ydist = self.opp_yplus / self.re_tau
        c.rk_step = self.rk_step
        c.scalar_bc = self.scalar_bc
        return c

    def bulk_u_value(self) -> float:
        return self.bulk_u if self.bulk_u is not None else 1.0


def preset(name: str) -> ChannelConfig:
    """The workloads of the benchmark description, by name."""
    if name == "re180_64":      # small, CPU-runnable; opposition control off
        return ChannelConfig(64, 65, 64, re_tau=180.0, lmfx=1.0, lmfz=0.5, dt=0.005, name=name)
    if name == "re180_192":     # production resolution, opposition control at y+ = 10, mixed-type scalar BC
        return ChannelConfig(192, 129, 160, re_tau=180.0, lmfx=2.0, lmfz=2.0 / 3.0, dt=0.0025,
                             opp_alpha=1.0, opp_yplus=10.0, scalar_bc=PB_SCALAR_MIXED, name=name)
    if name == "re1000":        # large single-GPU share; opposition control at y+ = 15
        return ChannelConfig(1024, 385, 768, re_tau=1000.0, lmfx=1.0, lmfz=0.5, dt=4.0e-4,
                             opp_alpha=1.0, opp_yplus=15.0, name=name)
    if name == "tiny":          # unit-test size
        return ChannelConfig(16, 33, 12, re_tau=180.0, lmfx=1.0, lmfz=0.5, dt=0.005, name=name)
    raise KeyError(name)


def physical_wavenumbers(cfg: ChannelConfig):
    """kx, kz index tables (module_ppdfftw.f90:347-384) times 1/LMF."""
    import numpy as np
    nh = cfg.nz // 2
    kx = np.arange(cfg.nxh, dtype=np.int64)
    kz = np.empty(cfg.nz, dtype=np.int64)
    kz[: nh + 1] = np.arange(nh + 1)
    kz[nh + 1:] = np.arange(nh + 2, cfg.nz + 1) - 2 * nh - 1
    return kx, kz, kx / cfg.lmfx, kz / cfg.lmfz


# grids of the ensemble workload (Re_tau = 180), with the box and a CFL-stable dt
ENSEMBLE_GRIDS = (
    (64, 65, 64, 1.0, 0.5, 0.005),
    (96, 97, 96, 1.0, 0.5, 0.0033),
    (128, 97, 128, 1.0, 0.5, 0.0025),
    (192, 129, 160, 2.0, 2.0 / 3.0, 0.0025),
)


ENSEMBLE_ALPHA_MAX = 1.0


def ensemble_configs(n: int = 24, seed: int = 0xC0FFEE + 3) -> list:
    """The ensemble workload: n independent Re_tau = 180 channels, grid drawn
    from ENSEMBLE_GRIDS, opposition gain uniform in [0, ENSEMBLE_ALPHA_MAX],
    detection plane y+ in {5, 10, 15, 20}, Prandtl number in {0.71, 1.0, 2.0},
    all from one seeded generator.  The gain stops at 1: the wall value
    v(+-1) = -alpha v(y_d) is set from the substep's input, so a near-wall
    mode is multiplied by about -alpha every substep, and gains above 1
    diverge (alpha = 1.26 at y+ = 5 and 20 on 64 x 65 x 64: the CFL number
    grows from 0.65 to 2.6 within 12 steps).  The time step follows the
    reference's CFL rule every step (dt_factor = 0, CFL = 0.8,
    dft_x_realspace.f90:489-498): the wall transpiration of the control grows
    as the perturbation develops, and |v(wall)| dyr(wall) with the fine
    near-wall spacing raises the CFL number (with a fixed dt, alpha = 0.25 at
    y+ = 20 goes from 0.62 to 1.03 in 150 steps, the uncontrolled flow stays
    at 0.62-0.68 over 300)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny, nz, lmfx, lmfz, dt = ENSEMBLE_GRIDS[int(rng.integers(len(ENSEMBLE_GRIDS)))]
        alpha = float(rng.uniform(0.0, ENSEMBLE_ALPHA_MAX))
        yplus = float(rng.choice([5.0, 10.0, 15.0, 20.0]))
        pr = float(rng.choice([0.71, 1.0, 2.0]))
        out.append(ChannelConfig(nx, ny, nz, re_tau=180.0, lmfx=lmfx, lmfz=lmfz, dt=dt, pr=pr, opp_alpha=alpha,
                                 opp_yplus=yplus, dt_factor=0.0, cfl=0.8, name=f"ens{i}_{nx}x{ny}x{nz}"))
    return out


def roofline_bytes_per_step(cfg: ChannelConfig) -> int:
    """Algorithmic HBM bytes of one RK3 step: 3 substeps x 30 spectral
    field passes x 16 B x ny*(nx/2)*nz."""
    return 3 * 30 * 16 * cfg.n_spec


PI = math.pi
