"""Benchmark: FP64 RK3-CN steps/s of the channel time step on one MI355X.

Workload: Re_tau = 1000 channel, 2pi x 2 x pi, 1024 x 385 x 768 (order-8
B-splines on tanh-stretched knots, stretch 2.0), passive scalar Pr = 0.71,
opposition control alpha = 1 at y+ = 15, synthetic seeded turbulent-mean
plus divergence-free perturbation state (no dataset).  One "step" is one
full RK3 step (3 substeps) of the whole field, state resident in HBM.

Prints ONE JSON line.  Device time of every stage is measured live with HIP
events on the library's stream; the dominant kernel's algorithmic bytes per
launch divided by its mean event time gives roofline.achieved.  The CPU
baseline is the C++/OpenMP oracle (the reference algorithm restated) timed
on a bounded sample of the same workload and extrapolated to one step.
Beside the headline the line carries three objects of other workloads: the
gradient of the discrete adjoint against its forward window (`adjoint`),
2000 steps of the 192 x 129 x 160 production case with statistics every 10
steps, statistics steps timed apart (`statistics`), and the 24-channel
ensemble against its channels run one after another (`ensemble`).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFL_LIMIT = 1.5  # largest pi*dt*velmax timed (RK3's imaginary-axis limit is sqrt(3))
STAGES = ["oppvd", "zinv", "xphys", "velmax", "zfwd", "wavespace"]
# stages that are a single kernel each (the wall-normal stage is a chain of
# sweep kernels on three streams); the roofline object names the single
# kernel with the largest time per step
KERNEL_OF_STAGE = {"oppvd": "opp_kernel", "zinv": "zinv_stream_kernel", "xphys": "xphys_stream_kernel",
                   "zfwd": "zfwd_stream_kernel"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "re1000_1024x385x768_hbm_traffic.txt")


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (already scaled by the copy calibration),
    or None when the summary is absent."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            for line in f:
                parts = line.split()
                if parts and parts[0].startswith(kernel) and len(parts) >= 6:
                    return (float(parts[-4]) + float(parts[-3])) * 1e9
    except (OSError, ValueError):
        pass
    return None


def cfl_number(dt, outs):
    """pi * dt * velmax of a step (velmax is the CFL maximum of its first
    substep's input, dft_x_realspace.f90:125-127); nan if not finite."""
    v = outs[0]["velmax"]
    return math.pi * dt * v if math.isfinite(v) else float("nan")


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def algorithmic_bytes(cfg):
    """Minimal HBM bytes per launch of each stage (inputs read once, outputs
    written once), in units of one spectral field pass = 16 B * ny*(nx/2)*nz.
    Padded z-pencil fields are 1.5 passes each."""
    f = 16.0 * cfg.n_spec
    return {
        "oppvd": 1.0 * f,                  # read v
        "zinv": (4 + 4 * 1.5) * f,         # read u,v,w,theta; write 4 padded
        "xphys": (4 * 1.5 + 8 * 1.5) * f,  # read 4 padded; write 8 padded
        "velmax": 0.0,
        "zfwd": (8 * 1.5 + 8) * f,         # read 8 padded; write 8 spectra
        "wavespace": (8 + 3 + 4 + 3) * f,  # read 8 products + 3 stores; write u,v,w,theta + 3 stores
    }


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain copy of known size for counter calibration: each dispatch
    reads and writes exactly nbytes."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    if hip.hipMemset(a, 0, ctypes.c_size_t(nbytes)) != 0:
        raise RuntimeError("hipMemset failed")
    for _ in range(reps):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)
    print(f"[bench] copy calibration: {reps} copies of {nbytes} bytes", file=sys.stderr, flush=True)


def ensemble_leg(steps, warmup, reps=3):
    """The ensemble workload (24 uneven Re_tau = 180 channels, seeded gains,
    detection planes and Prandtl numbers): every channel stepped with pb_step
    one after another (the plain path), against the tuned ensemble, one
    captured graph per ensemble step.  Both run warmup + steps steps from the
    same states; the final states must be bit-identical."""
    import collections
    import numpy as np
    from channel_passive_amd import capi
    from channel_passive_amd.config import ensemble_configs
    from channel_passive_amd.initial import make_state

    cfgs = ensemble_configs(24)
    # every channel takes its dt from the CFL rule on substep 1 (dt_factor =
    # 0); the values passed in are only starting values
    dts = [c.dt for c in cfgs]
    ens = capi.Ensemble(cfgs)
    tuned = ens.tune(reps=reps)
    plain = []
    for i, c in enumerate(cfgs):
        S = make_state(c, seed=0xC0FFEE + 3 + i)
        ens.channel(i).set_state(S)
        ens.channel(i).init_stores()
        ch = capi.Channel(c)
        ch.set_state(S)
        ch.init_stores()
        plain.append(ch)
    print(f"[bench] ensemble: {len(cfgs)} channels set, tuned step {tuned[-1]:.3f} ms", file=sys.stderr, flush=True)
    for _ in range(warmup):
        ens.step(dts)
        for c, ch in zip(cfgs, plain):
            ch.step(c.dt)
    for ch in plain:
        ch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for c, ch in zip(cfgs, plain):
            ch.step(c.dt)
    t_seq = time.perf_counter() - t0
    cfl_max = 0.0
    print(f"[bench] ensemble: sequential plain {1e3 * t_seq / steps:.3f} ms per step", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        outs = ens.step(dts)
    t_ens = time.perf_counter() - t0
    for c, o in zip(cfgs, outs):
        cfl_max = max(cfl_max, math.pi * o[0]["dt"] * o[0]["velmax"])
    print(f"[bench] ensemble: one graph {1e3 * t_ens / steps:.3f} ms per step", file=sys.stderr, flush=True)
    same = True
    for i, ch in enumerate(plain):
        Sa, Ra = ens.channel(i).get_state()
        Sb, Rb = ch.get_state()
        same = same and bool(np.array_equal(Sa, Sb) and np.array_equal(Ra, Rb))
        ch.close()
    ens.close()
    grids = collections.Counter(f"{c.nx}x{c.ny}x{c.nz}" for c in cfgs)
    return {"channels": len(cfgs), "grids": dict(grids), "steps": steps, "warmup": warmup,
            "sequential_ms_per_step": 1e3 * t_seq / steps, "ensemble_ms_per_step": 1e3 * t_ens / steps,
            "speedup": t_seq / t_ens, "bitwise_equal_to_plain": same, "dt": "CFL rule, CFL = 0.8",
            "cfl_number_last": round(cfl_max, 4),
            "tuned_channel_ms": [round(float(x), 4) for x in tuned[:-1]], "tuned_ensemble_ms": float(tuned[-1])}


def adjoint_leg(steps, ckpt_every=20, warmup=5):
    """The differentiation workload (Re_tau = 180, 192x129x160, opposition
    gain 0.8 at y+ = 10, a seeded Gaussian wall actuation of std 1e-3, the
    mixed-type scalar condition): J = window mean of tau_w + 0.5 q_w over
    `steps` steps and its gradient with respect to the gain and the
    actuation field (pb_gradient: forward with checkpoints every ckpt_every
    steps, recomputation, the transposed substeps), timed against the same
    window stepped forward alone."""
    import dataclasses
    import numpy as np
    from channel_passive_amd import capi
    from channel_passive_amd.config import preset
    from channel_passive_amd.initial import make_state

    cfg = dataclasses.replace(preset("re180_192"), opp_alpha=0.8, opp_yplus=10.0)
    ch = capi.Channel(cfg)
    rng = np.random.default_rng(0xC0FFEE + 4)
    act = 1e-3 * (rng.standard_normal((cfg.nz, cfg.nxh)) + 1j * rng.standard_normal((cfg.nz, cfg.nxh)))
    act[0, 0] = 0.0
    ch.set_control(cfg.opp_alpha, act)
    ch.set_state(make_state(cfg, seed=0xC0FFEE + 4))
    ch.init_stores()
    for _ in range(warmup):
        ch.step(cfg.dt)
    S0, RK0 = ch.get_state()
    ch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ch.step(cfg.dt)
    ch.sync()
    t_fwd = time.perf_counter() - t0
    ch.set_state(S0, RK0)
    ch.gradient(2, cfg.dt, ckpt_every=1)  # allocations and first launches outside the timing
    ch.set_state(S0, RK0)
    ch.sync()
    t0 = time.perf_counter()
    J, dJa, g = ch.gradient(steps, cfg.dt, ckpt_every=ckpt_every)
    t_grad = time.perf_counter() - t0
    ch.close()
    print(f"[bench] adjoint: forward {1e3 * t_fwd / steps:.3f} ms/step, gradient {1e3 * t_grad / steps:.3f} ms/step",
          file=sys.stderr, flush=True)
    return {"workload": f"Re_tau=180 {cfg.nx}x{cfg.ny}x{cfg.nz}, alpha={cfg.opp_alpha} at y+={cfg.opp_yplus:g}, "
                        "wall actuation std 1e-3, mixed-type scalar condition",
            "window_steps": steps, "ckpt_every": ckpt_every, "dt": cfg.dt,
            "forward_ms_per_step": 1e3 * t_fwd / steps, "gradient_ms_per_step": 1e3 * t_grad / steps,
            "cost_ratio": t_grad / t_fwd, "J": J, "dJ_dalpha": dJa, "dJ_da_max_abs": float(np.abs(g).max())}


def statistics_leg(steps=2000, every=10, warmup=10):
    """The production-resolution workload with statistics (Re_tau = 180,
    192x129x160, opposition control alpha = 1 at y+ = 10, mixed-type scalar
    condition, dt from the CFL rule at CFL 0.8): `steps` steps, every
    `every`-th one a statistics step -- substep 1 with the pressure and
    get_statistics_2 (pb_substep_stats2), substeps 2 and 3, then
    get_statistics_1's profiles (pb_statistics), results copied to the
    host -- timed apart from the plain steps."""
    import dataclasses
    import numpy as np
    from channel_passive_amd import capi
    from channel_passive_amd import observables as ob
    from channel_passive_amd.config import preset
    from channel_passive_amd.initial import make_state

    cfg = dataclasses.replace(preset("re180_192"), dt_factor=0.0, cfl=0.8)
    ch = capi.Channel(cfg)
    ch.set_state(make_state(cfg, seed=0xC0FFEE + 1))
    ch.init_stores()
    quad = ob.Quadrature(ch.labs(), ch.bw)
    dt = cfg.dt
    for _ in range(warmup):
        dt = ch.step(dt)[-1]["dt"]
    t_plain = t_stats = 0.0
    n_stats, cfl_max, last = 0, 0.0, None
    for k in range(1, steps + 1):
        t0 = time.perf_counter()
        if k % every == 0:
            o1, _, _, _ = ch.substep_stats2(dt)
            o2 = ch.substep(2, o1["dt"])
            o3 = ch.substep(3, o1["dt"])
            prof = ch.statistics()
            t_stats += time.perf_counter() - t0
            n_stats += 1
            outs = [o1, o2, o3]
            last = (prof, o3)
        else:
            outs = ch.step(dt)
            t_plain += time.perf_counter() - t0
        dt = outs[-1]["dt"]
        cfl_max = max(cfl_max, math.pi * outs[0]["dt"] * outs[0]["velmax"])
    ch.close()
    prof, o3 = last
    n_plain = steps - n_stats
    tm = ob.theta_mean(prof, quad, "mixed", cfg.u_gal)
    print(f"[bench] statistics: plain {1e3 * t_plain / n_plain:.3f} ms/step, statistics step "
          f"{1e3 * t_stats / n_stats:.3f} ms", file=sys.stderr, flush=True)
    return {"workload": f"Re_tau=180 {cfg.nx}x{cfg.ny}x{cfg.nz}, alpha={cfg.opp_alpha} at y+={cfg.opp_yplus:g}, "
                        "mixed-type scalar condition, dt from the CFL rule (CFL 0.8)",
            "steps": steps, "statistics_every": every, "warmup": warmup,
            "plain_ms_per_step": 1e3 * t_plain / n_plain, "statistics_step_ms": 1e3 * t_stats / n_stats,
            "statistics_overhead_ms": 1e3 * (t_stats / n_stats - t_plain / n_plain),
            "total_s": t_plain + t_stats, "cfl_number_max": round(cfl_max, 4),
            "last_Cf": float(ob.skin_friction(o3["tau_w"], cfg.bulk_u_value())),
            "last_Nu": float(ob.nusselt(o3["q_w"], cfg.nu, cfg.pr, tm))}


ENSEMBLE_HW_QUEUES = "16"


def ensemble_subprocess(steps):
    """The ensemble comparison in a child process with GPU_MAX_HW_QUEUES = 16
    (the HIP runtime reads it at start-up; 24 channels on the default 4
    hardware queues overlap only four at a time).  Failures are reported in
    the object, the headline stands."""
    import subprocess
    env = dict(os.environ, GPU_MAX_HW_QUEUES=ENSEMBLE_HW_QUEUES)
    cmd = [sys.executable, os.path.abspath(__file__), "--ensemble-only", "--ensemble-steps", str(steps)]
    try:
        res = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=900)
        sys.stderr.write(res.stderr[-2000:])
        if res.returncode != 0:
            return {"error": f"exit {res.returncode}", "stderr_tail": res.stderr[-500:]}
        obj = json.loads(res.stdout.strip().splitlines()[-1])
        obj["gpu_max_hw_queues"] = int(ENSEMBLE_HW_QUEUES)
        return obj
    except Exception as exc:  # reported, not fatal
        return {"error": f"{type(exc).__name__}: {exc}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--config", default="re1000")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 24-channel ensemble comparison")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the gradient (discrete adjoint) leg")
    ap.add_argument("--adjoint-steps", type=int, default=400)
    ap.add_argument("--adjoint-only", action="store_true", help="run only the gradient leg and print its JSON")
    ap.add_argument("--no-statistics", action="store_true",
                    help="skip the 2000-step run with statistics every 10 steps (192x129x160)")
    ap.add_argument("--statistics-only", action="store_true",
                    help="run only the statistics leg and print its JSON")
    ap.add_argument("--statistics-steps", type=int, default=2000)
    ap.add_argument("--ensemble-only", action="store_true",
                    help="run only the ensemble comparison and print its JSON object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy 3 times (the reference kernel for "
                         "scaling rocprofv3 FETCH_SIZE/WRITE_SIZE to bytes) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.ensemble_only:
        print(json.dumps(ensemble_leg(args.ensemble_steps, 10)))
        return
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(args.adjoint_steps)))
        return
    if args.statistics_only:
        print(json.dumps(statistics_leg(args.statistics_steps)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the library runs on one GPU "
              "and has no multi-GPU (sharded) path", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import numpy as np
    from channel_passive_amd import capi
    from channel_passive_amd.config import preset, roofline_bytes_per_step
    from channel_passive_amd.initial import make_state

    cfg = preset(args.config)
    print(f"[bench] config {cfg.name}: {cfg.nx}x{cfg.ny}x{cfg.nz}", file=sys.stderr, flush=True)
    ch = capi.Channel(cfg)
    S = make_state(cfg, seed=0xC0FFEE + 2)
    ch.set_state(S)
    del S
    ch.init_stores()
    print(f"[bench] state set, device bytes {ch.device_bytes/1e9:.1f} GB", file=sys.stderr, flush=True)
    cfl_seen = 0.0
    for w in range(args.warmup):
        out = ch.step(cfg.dt)
        cfl_seen = max(cfl_seen, cfl_number(cfg.dt, out))
        if (w + 1) % 10 == 0 or w + 1 == args.warmup:
            print(f"[bench] warmup {w + 1}/{args.warmup}: pi*dt*velmax = {cfl_number(cfg.dt, out):.3f}",
                  file=sys.stderr, flush=True)
    if not cfl_seen <= CFL_LIMIT:
        raise SystemExit(f"refusing to time an unstable run: pi*dt*velmax = {cfl_seen:.3f} > {CFL_LIMIT} "
                         f"(RK3 is stable up to sqrt(3) on the imaginary axis)")
    L = capi.load()
    capi._check(L.pb_timing(ch.h, 1))
    ch.sync()
    t0 = time.perf_counter()
    last = None
    cfl_timed = 0.0
    for k in range(args.steps):
        last = ch.step(cfg.dt)  # raises on a non-finite state (PB_E_NONFINITE)
        cfl_timed = max(cfl_timed, last[0]["velmax"])
    ch.sync()
    t1 = time.perf_counter()
    cfl_timed = math.pi * cfg.dt * cfl_timed
    ms = (ctypes.c_double * 6)()
    cnt = (ctypes.c_int64 * 6)()
    capi._check(L.pb_timing_read(ch.h, ms, cnt))
    capi._check(L.pb_timing(ch.h, 0))
    elapsed = t1 - t0
    steps_per_s = args.steps / elapsed
    stage = {STAGES[i]: {"ms_total": ms[i], "launches": int(cnt[i]),
                         "ms_mean": (ms[i] / cnt[i]) if cnt[i] else 0.0} for i in range(6)}
    alg = algorithmic_bytes(cfg)
    dom = max(KERNEL_OF_STAGE, key=lambda s: stage[s]["ms_total"])
    achieved = alg[dom] / (stage[dom]["ms_mean"] * 1e-3) / 1e9
    traffic = measured_traffic(KERNEL_OF_STAGE[dom]) if cfg.name == "re1000" else None
    stage_frac = {s: round(alg[s] / (stage[s]["ms_mean"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                  for s in STAGES if alg[s] > 0 and stage[s]["ms_mean"] > 0}
    step_bytes = roofline_bytes_per_step(cfg)
    print(f"[bench] {args.steps} steps in {elapsed:.3f} s; stages: " +
          ", ".join(f"{s} {stage[s]['ms_mean']:.3f} ms" for s in STAGES), file=sys.stderr, flush=True)
    # every mode of the final state enters the Parseval sums of the
    # statistics profiles: a NaN or Inf anywhere makes them non-finite
    stats = ch.statistics()
    if not all(np.all(np.isfinite(v)) for v in stats.values()):
        raise SystemExit("the state after the timed steps is not finite")
    if not cfl_timed <= CFL_LIMIT:
        raise SystemExit(f"the timed run went CFL-unstable: pi*dt*velmax = {cfl_timed:.3f}")
    ch.close()

    cpu = None
    if not args.no_cpu_baseline:
        from oracle.oracle import sample_times
        nplanes, nlines = 96, 131072
        tn, tw, nt = sample_times(cfg, nplanes, nlines)
        t_step = 3 * (tn * cfg.ny / nplanes + tw * (cfg.nxh * cfg.nz) / nlines)
        cpu = {"value": 1.0 / t_step, "unit": "steps/s", "cores": nt, "cpu_model": cpu_model(), "kind": "port",
               "sample": f"oracle (reference algorithm restated, C++/OpenMP, own FFT) at the full grid: "
                         f"nonlinear stage on {nplanes} of {cfg.ny} y-planes ({tn:.2f} s) and wavespace on "
                         f"{nlines} of {cfg.nxh*cfg.nz} (kx,kz) systems ({tw:.2f} s), extrapolated to 3 substeps"}

    out = {
        "metric": "FP64 RK3 steps/s at Re_tau 1000",
        "value": steps_per_s,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded turbulent-mean profile + divergence-free perturbation)",
        "config": {"workload": f"channel Re_tau={cfg.re_tau:g} {cfg.nx}x{cfg.ny}x{cfg.nz}, opposition control "
                               f"alpha={cfg.opp_alpha} at y+={cfg.opp_yplus:g}, Pr={cfg.pr}",
                   "nx": cfg.nx, "ny": cfg.ny, "nz": cfg.nz, "dt": cfg.dt, "nu": cfg.nu,
                   "units": "bulk (U_b = 1, h = 1)",
                   "cfl_number_max": round(max(cfl_seen, cfl_timed), 4)},
        "roofline": {"bound": "hbm", "kernel": KERNEL_OF_STAGE[dom], "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes": alg[dom],
                     "step_frac": step_bytes * steps_per_s / 1e9 / HBM_PEAK_GBS},
        "stages_ms": {s: round(stage[s]["ms_mean"], 4) for s in STAGES},
        "stages_roofline_frac": stage_frac,
        "cpu_baseline": cpu,
    }
    if not args.no_adjoint:
        try:
            out["adjoint"] = adjoint_leg(args.adjoint_steps)
        except Exception as exc:  # reported, not fatal: the headline stands
            out["adjoint"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_statistics:
        try:
            out["statistics"] = statistics_leg(args.statistics_steps)
        except Exception as exc:  # reported, not fatal: the headline stands
            out["statistics"] = {"error": f"{type(exc).__name__}: {exc}"}
    if not args.no_ensemble:
        out["ensemble"] = ensemble_subprocess(args.ensemble_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
