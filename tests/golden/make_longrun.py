"""Long-run statistics of the oracle (the CPU restatement of the reference's
algorithm) for the benchmark's small case: Re_tau = 180, 64x65x64, dt =
0.005 (bulk units), the seeded perturbed laminar state of make_state (seed
SEED), NSTEPS steps, get_statistics_1's profiles every EVERY steps; the mean
and rms profiles, Cf and Nu averaged over the samples of the second half.
Writes longrun_re180_64.npz beside this file (the GPU long-run test compares
the device's run of the same case against it).

Run from the repository root: python tests/golden/make_longrun.py
(about 20 minutes on 8 cores).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from channel_passive_amd import observables as ob  # noqa: E402
from channel_passive_amd.config import preset  # noqa: E402
from channel_passive_amd.initial import make_state  # noqa: E402
from oracle import oracle as O  # noqa: E402

NSTEPS, EVERY, SEED = 2000, 10, 0xC0FFEE
OUT = os.path.join(HERE, "longrun_re180_64.npz")


def sample(cfg, st, out, quad):
    """The quantities averaged: U, Theta, the four rms profiles, Cf, Nu."""
    rms = ob.rms_profiles(st, cfg.u_gal)
    tm = ob.theta_mean(st, quad, "bulk", cfg.u_gal)
    return {"U": np.asarray(st["U"], float) - cfg.u_gal, "Theta": np.asarray(st["Theta"], float),
            "u_rms": rms["u_rms"], "v_rms": rms["v_rms"], "w_rms": rms["w_rms"], "t_rms": rms["t_rms"],
            "Cf": ob.skin_friction(out["tau_w"], cfg.bulk_u_value()),
            "Nu": ob.nusselt(out["q_w"], cfg.nu, cfg.pr, tm)}


def main():
    O.build()
    cfg = preset("re180_64")
    orc = O.Oracle(cfg)
    quad = ob.Quadrature({"d0": orc.d0, "integ": orc.integ}, orc.bw)
    S = make_state(cfg, seed=SEED)
    RK = orc.init_stores(S)
    series, t0 = [], time.time()
    for k in range(1, NSTEPS + 1):
        outs = orc.step(S, RK, cfg.dt)
        if k % EVERY == 0:
            series.append((k, sample(cfg, orc.statistics(S), outs[-1], quad)))
        if k % 100 == 0:
            print(f"step {k}/{NSTEPS} ({time.time() - t0:.0f} s)", flush=True)
    half = [s for k, s in series if k > NSTEPS // 2]
    avg = {key: np.mean([s[key] for s in half], axis=0) for key in half[0]}
    np.savez(OUT, nsteps=NSTEPS, every=EVERY, seed=SEED, steps=np.array([k for k, _ in series]),
             Cf_series=np.array([s["Cf"] for _, s in series]), Nu_series=np.array([s["Nu"] for _, s in series]),
             **{"avg_" + k: v for k, v in avg.items()})
    print("wrote", OUT, {k: (float(v) if np.ndim(v) == 0 else float(np.max(v))) for k, v in avg.items()})


if __name__ == "__main__":
    main()
