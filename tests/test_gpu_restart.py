"""GPU: the restart kinds of the reference (from v/g and from g/phi) against
the oracle's restatement, bit for bit; the checkpoint export of data_store's
datasets after a step, bit for bit against the oracle's state and stores;
a restart from the checkpoint continues the run identically."""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import preset
from channel_passive_amd.initial import make_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from channel_passive_amd import capi as C
    return C


@pytest.mark.parametrize("name", ["re180_64", "tiny_opp"])
def test_restart_kinds_bitwise(capi, oracle_lib, name):
    cfg = dataclasses.replace(preset("re180_64") if name == "re180_64" else preset("tiny"), u_gal=0.1)
    if name == "tiny_opp":
        cfg = dataclasses.replace(cfg, opp_alpha=0.5, opp_yplus=20.0)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=3)
    RK = orc.init_stores(S)
    orc.step(S, RK, cfg.dt)
    means = (S[0, 0, 0].real + cfg.u_gal, S[2, 0, 0].real, S[3, 0, 0].real)
    S_vg, RK_vg = orc.init_from_vg(S[1], RK[0], S[3], *means)
    S_gp, RK_gp = orc.init_from_gphi(RK[0], RK[1], S[3], *means)
    ch = capi.Channel(cfg)
    ch.set_state_vg(S[1], RK[0], S[3], *means)
    Sg, RKg = ch.get_state()
    assert np.array_equal(Sg, S_vg) and np.array_equal(RKg, RK_vg)
    ch.set_state_gphi(RK[0], RK[1], S[3], *means)
    Sg, RKg = ch.get_state()
    assert np.array_equal(Sg, S_gp) and np.array_equal(RKg, RK_gp)
    ch.close()


def test_checkpoint_and_continue(capi, oracle_lib):
    cfg = dataclasses.replace(preset("re180_64"), u_gal=0.1)
    orc = oracle_lib.Oracle(cfg)
    S = make_state(cfg, seed=4)
    a = capi.Channel(cfg)
    a.set_state(S)
    a.init_stores()
    for _ in range(2):
        a.step(cfg.dt)
    cp = a.checkpoint()
    Sa, RKa = a.get_state()
    assert np.array_equal(cp["u"], Sa[0]) and np.array_equal(cp["v"], Sa[1]) and np.array_equal(cp["w"], Sa[2])
    assert np.array_equal(cp["theta"], Sa[3]) and np.array_equal(cp["g"], RKa[0]) and np.array_equal(cp["phi"], RKa[1])
    assert np.array_equal(cp["U_mean"], Sa[0, 0, 0].real + cfg.u_gal) and np.array_equal(cp["T_mean"], Sa[3, 0, 0].real)
    # restart from the checkpoint (g/phi kind) and continue beside the original
    b = capi.Channel(cfg)
    b.set_state_gphi(cp["g"], cp["phi"], cp["theta"], cp["U_mean"], cp["W_mean"], cp["T_mean"])
    Sb, _ = b.get_state()
    for f in range(4):
        assert np.linalg.norm(Sb[f] - Sa[f]) <= 1e-11 * np.linalg.norm(Sa[f]), f
    ob, oa = b.step(cfg.dt), a.step(cfg.dt)
    Sa, _ = a.get_state()
    Sb, _ = b.get_state()
    for f in range(4):
        assert np.linalg.norm(Sb[f] - Sa[f]) <= 1e-11 * np.linalg.norm(Sa[f]), f
    a.close()
    b.close()
