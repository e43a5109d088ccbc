"""The workloads' units and time steps are CFL-stable.

The reference picks dt from dt <= CFL / (pi * velmax) (read_input_file.f90:107,
dft_x_realspace.f90:489-498) with velmax = max(|u| dxr + |v| dyr + |w| dzr)
over the padded physical grid.  RK3 is stable on the imaginary axis up to
sqrt(3), so every preset's fixed dt must keep pi * dt * velmax well below it
on its seeded initial state, and the small workload must run its full 100
steps bounded.
"""
import math

import numpy as np
import pytest

from channel_passive_amd.config import dean_bulk_reynolds, preset
from channel_passive_amd.initial import collocation_points, make_state

CFL_MAX = 1.0   # pi * dt * velmax of each preset's initial state (limit sqrt(3) ~ 1.73)


def velmax_numpy(cfg, S, nx_s, nz_s):
    """The CFL maximum of a state whose modes all sit below nx_s/2, nz_s/2,
    sampled on an nx_s x nz_s physical grid with numpy's inverse FFT (an
    independent restatement of dft_x_realspace.f90:125-127), using the full
    grid's dxr, dzr and dyr."""
    y = collocation_points(cfg)
    d = np.empty(cfg.ny)
    d[0], d[-1] = abs(y[1] - y[0]), abs(y[-1] - y[-2])
    d[1:-1] = np.minimum(np.abs(y[1:-1] - y[:-2]), np.abs(y[1:-1] - y[2:]))
    dyr = 1.0 / (cfg.dy_factor * d)
    dxr = cfg.nx / (2 * math.pi * cfg.lmfx)
    dzr = cfg.nz / (2 * math.pi * cfg.lmfz)
    nz_st, nxh_st = S.shape[1], S.shape[2]   # the state's own grid (may be smaller)
    nh = nz_st // 2
    kzi = np.concatenate([np.arange(nh + 1), np.arange(nh + 2, nz_st + 1) - 2 * nh - 1])
    phys = []
    for f in range(3):
        full = np.zeros((nz_s, nx_s, cfg.ny), complex)
        for j in range(nz_st):
            for ix in range(nxh_st):
                a = S[f, j, ix]
                if not np.any(a):
                    continue
                assert abs(kzi[j]) < nz_s // 2 and ix < nx_s // 2, "mode outside the sample grid"
                full[kzi[j] % nz_s, ix] += a
                if ix > 0:
                    full[(-kzi[j]) % nz_s, nx_s - ix] += np.conj(a)
        phys.append(np.real(np.fft.ifft2(full, axes=(0, 1))) * nx_s * nz_s)
    vel = np.abs(phys[0]) * dxr + np.abs(phys[1]) * dyr + np.abs(phys[2]) * dzr
    return float(vel.max())


def test_bulk_units():
    """U_b = 1 and nu = 1/Re_b with Dean's Re_tau = 0.09 (2 Re_b)^0.88."""
    cfg = preset("re180_64")
    assert cfg.bulk_u_value() == 1.0
    re_b = dean_bulk_reynolds(180.0)
    assert abs(0.09 * (2 * re_b) ** 0.88 - 180.0) < 1e-9
    assert abs(cfg.nu * re_b - 1.0) < 1e-15
    # the laminar start has bulk velocity 1 (initial_step.f90:83)
    S = make_state(cfg, amp=0.0)
    y = collocation_points(cfg)
    assert np.allclose(S[0, 0, 0].real, 1.5 * (1 - y * y))


@pytest.mark.parametrize("name", ["tiny", "re180_64", "re180_192"])
def test_preset_cfl_oracle(oracle_lib, name):
    """Oracle velmax of the seeded initial state (the bench seed and the
    test seed) against the preset's dt; the numpy restatement agrees."""
    cfg = preset(name)
    orc = oracle_lib.Oracle(cfg)
    for seed in (0xC0FFEE, 0xC0FFEE + 2):
        S = make_state(cfg, seed=seed)
        _, vm = orc.nonlinear(S)
        cfl = math.pi * cfg.dt * vm
        assert 0.0 < cfl <= CFL_MAX, (name, seed, cfl)
        # brute force on the same padded 3nx/2 x 3nz/2 grid
        vn = velmax_numpy(cfg, S, 3 * cfg.nx // 2, 3 * cfg.nz // 2)
        assert abs(vn - vm) <= 1e-12 * vm, (vn, vm)


def test_preset_cfl_re1000():
    """The 1024 x 385 x 768 workload: its perturbation holds modes |k| <= 8
    only, so a 64 x 48 sample of the same field (same collocation points,
    same dxr, dzr, dyr of the full grid) gives its CFL maximum."""
    cfg = preset("re1000")
    small = type(cfg)(**{**cfg.__dict__, "nx": 64, "nz": 48})
    S = make_state(small, seed=0xC0FFEE + 2)
    vm = velmax_numpy(cfg, S, 128, 96)
    cfl = math.pi * cfg.dt * vm
    assert 0.3 <= cfl <= CFL_MAX, cfl


def test_re180_64_hundred_steps(oracle_lib):
    """The small workload's full 100 RK3-CN steps on the oracle at dt = 0.005:
    finite, flow rate held, CFL number and perturbation bounded."""
    cfg = preset("re180_64")
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg)
    RK = orc.init_stores(S)
    v0 = np.abs(S[1]).max()
    vm = []
    for _ in range(100):
        outs = orc.step(S, RK, cfg.dt)
        vm.append(outs[0]["velmax"])
        assert all(math.isfinite(v) for o in outs for v in o.values())
    assert np.isfinite(S).all() and np.isfinite(RK).all()
    cfl = math.pi * cfg.dt * max(vm)
    assert cfl <= CFL_MAX, cfl
    assert max(vm) <= 1.25 * vm[0]
    assert np.abs(S[1]).max() <= 3.0 * v0
    c = oracle_lib.band_solve(orc.d0lu, orc.bw, S[0, 0, 0].copy())
    assert abs(0.5 * orc.integ @ c.real - 1.0) <= 1e-12


def test_oracle_adaptive_dt_threads_through(oracle_lib):
    """dt_factor = 0: substep 1 sets dt = CFL / (pi * velmax)
    (dft_x_realspace.f90:489-498) and substeps 2 and 3 use that dt."""
    import dataclasses
    cfg = dataclasses.replace(preset("tiny"), dt_factor=0.0, cfl=0.8)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=9)
    RK = orc.init_stores(S)
    outs = orc.step(S, RK, 1.0)
    assert outs[0]["dt"] == outs[1]["dt"] == outs[2]["dt"]
    assert abs(outs[0]["dt"] - 0.8 / math.pi / outs[0]["velmax"]) <= 1e-15 * outs[0]["dt"]
    assert np.all(np.isfinite(S))
