"""GPU: the captured step (pb_step_graph) and the ensemble (pb_ensemble_*).

* pb_step_graph against pb_step on the same state: bit-identical states,
  stores and outputs after every step, with dt changing from step to step
  and eager work (state read-back) between the graph replays.
* an ensemble of uneven channels (four grids, opposition gains, detection
  planes, Prandtl numbers, both scalar conditions), tuned by the
  autotuner, against each channel stepped alone with pb_step on an untuned
  context: bit-identical after every step.
"""
import dataclasses

import numpy as np
import pytest

from channel_passive_amd.config import ensemble_configs, preset
from channel_passive_amd.initial import make_state

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def capi():
    from channel_passive_amd import capi as C
    return C


@pytest.mark.parametrize("name", ["re180_64", "opp_mixed"])
def test_step_graph_bitwise(capi, oracle_lib, name):
    cfg = preset("re180_64")
    if name == "opp_mixed":
        cfg = dataclasses.replace(cfg, opp_alpha=0.8, opp_yplus=15.0, scalar_bc=1, pr=2.0)
    S = make_state(cfg, seed=19)
    a, b = capi.Channel(cfg), capi.Channel(cfg)
    for ch in (a, b):
        ch.set_state(S)
        ch.init_stores()
    for k, f in enumerate((1.0, 0.9, 1.1, 1.0)):
        oa = a.step(cfg.dt * f)
        ob = b.step_graph(cfg.dt * f)
        Sa, RKa = a.get_state()
        Sb, RKb = b.get_state()
        assert np.array_equal(Sa, Sb) and np.array_equal(RKa, RKb), k
        assert oa == ob, k
    a.close()
    b.close()


def test_step_graph_refuses_timing(capi):
    cfg = preset("tiny")
    ch = capi.Channel(cfg)
    ch.set_state(make_state(cfg, seed=1))
    ch.init_stores()
    capi._check(capi.load().pb_timing(ch.h, 1))
    with pytest.raises(capi.PbError, match="timing"):
        ch.step_graph(cfg.dt)
    capi._check(capi.load().pb_timing(ch.h, 0))
    ch.step_graph(cfg.dt)
    ch.close()


def small_ensemble():
    cfgs = ensemble_configs(24)
    pick = {}
    for c in cfgs:  # one channel of each grid, as drawn
        pick.setdefault((c.nx, c.ny, c.nz), c)
    chosen = list(pick.values())
    chosen.append(dataclasses.replace(preset("tiny"), scalar_bc=1, pr=2.0, opp_alpha=0.3, opp_yplus=20.0))
    return chosen


def test_ensemble_bitwise_against_plain(capi):
    cfgs = small_ensemble()
    assert len({(c.nx, c.ny, c.nz) for c in cfgs}) >= 4
    ens = capi.Ensemble(cfgs)
    ms = ens.tune(reps=2)
    assert np.all(np.isfinite(ms)) and np.all(ms > 0)
    plain = []
    for i, cfg in enumerate(cfgs):
        S = make_state(cfg, seed=100 + i)
        ens.channel(i).set_state(S)
        ens.channel(i).init_stores()
        ch = capi.Channel(cfg)
        ch.set_state(S)
        ch.init_stores()
        plain.append(ch)
    for k in range(3):
        outs = ens.step([c.dt for c in cfgs])
        for i, (cfg, ch) in enumerate(zip(cfgs, plain)):
            op = ch.step(cfg.dt)
            assert outs[i] == op, (k, i)
            Se, RKe = ens.channel(i).get_state()
            Sp, RKp = ch.get_state()
            assert np.array_equal(Se, Sp) and np.array_equal(RKe, RKp), (k, i, cfg.name)
    for ch in plain:
        ch.close()
    ens.close()
