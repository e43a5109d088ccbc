"""Restart paths of the reference (initial_step.f90:20-45): from u, v, w
(set_initial_data_uvw, the stores), from v and g (set_initial_data_vg) and
from g and phi (set_initial_data_gphi), pinned on the oracle by round trips:
a divergence-free state with v = v' = 0 at the walls, its stores g, phi, its
theta and mean lines, restarted through either path, gives back u, v, w and
theta to rounding (the g/phi path solves a Poisson problem for v), and the
phi store of the v/g path equals set_initial_data_uvw's bit for bit."""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("name", ["tiny", "re180_64"])
def test_restart_round_trips(oracle_lib, name):
    cfg = dataclasses.replace(preset(name), u_gal=0.2)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=5)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)   # a developed state: v, v' = 0 at the walls, divergence free
    RK = orc.init_stores(S)   # its stores g = i kz u - i kx w, phi = Lap v
    means = (S[0, 0, 0].real + cfg.u_gal, S[2, 0, 0].real, S[3, 0, 0].real)
    S_vg, RK_vg = orc.init_from_vg(S[1], RK[0], S[3], *means)
    for f in range(4):
        assert rel(S_vg[f], S[f]) <= 1e-12, ("vg", f)
    assert np.array_equal(RK_vg[1][:, 1:], RK[1][:, 1:]) and np.array_equal(RK_vg[1][1:], RK[1][1:])
    S_gp, _ = orc.init_from_gphi(RK[0], RK[1], S[3], *means)
    for f in range(4):
        assert rel(S_gp[f], S[f]) <= 1e-11, ("gphi", f, rel(S_gp[f], S[f]))
    # the mean lines come back without the Galilean shift
    assert np.allclose(S_gp[0, 0, 0].real, means[0] - cfg.u_gal, rtol=0, atol=1e-15)


def test_gphi_restart_detects_a_wrong_phi(oracle_lib):
    """A phi that is not the Laplacian of the state's v gives a different v:
    the round trip is a real check of the Poisson solve."""
    cfg = preset("tiny")
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=6)
    RK = orc.init_stores(S)
    means = (S[0, 0, 0].real, S[2, 0, 0].real, S[3, 0, 0].real)
    S_gp, _ = orc.init_from_gphi(RK[0], 1.01 * RK[1], S[3], *means)
    assert rel(S_gp[1], S[1]) > 1e-3
