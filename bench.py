"""Benchmark of the Hotspice stepping path on one MI355X.

Default workload (the plain Neel configuration): in-plane pinwheel ASI 256x256 (32768
magnets), full 511x511 dipolar kernels + 1 mT Zeeman field, Neel-Arrhenius, FP32 + Kahan
fields with an FP64 recompute every 1024 switches. One benchmark step = 1024 Neel switch
attempts + the FP64 recompute, i.e. one recompute period; `value` = switch attempts/s.

Prints ONE JSON line (see the keys below). `roofline` is for the dominant kernel (the persistent
Neel step kernel, one launch per 1024 attempts): algorithmic bytes per launch (attempts per launch
x neel_bytes_per_attempt) / average launch duration, measured with HIP events recorded on the
launch stream around every launch in the timed region. `cpu_baseline` times the FP64 NumPy restatement of
the reference on one host core, on a bounded sample of the same workload.

    python bench.py [--steps K] [--warmup W] [--cpu-seconds S] [--no-cpu]
"""
import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PERIOD = 1024


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


def cpu_baseline(seconds):
    """FP64 NumPy restatement of the reference (oracle/model.py) on the same workload, one core,
    with NumPy generators as in the reference (no replay needed for timing)."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import model
    from hotspice_amd.workloads import pinwheel_neel_inputs
    p = pinwheel_neel_inputs()
    om = model.make_magnets("IP_Pinwheel", p["a"], n=p["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"],
                            pattern=p["pattern"], energies=[model.Dipolar(recompute="fft"), model.Zeeman(*p["field"])],
                            params=dict(UPDATE_SCHEME="Neel"), seed=1, rng="numpy")
    for _ in range(3):  # warm-up
        om.update(t_max=math.inf)
    a0 = om.attempted_switches
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        om.update(t_max=math.inf)
    dt = time.perf_counter() - t0
    n = om.attempted_switches - a0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return dict(value=n / dt, unit="switch attempts/s", cores=1, kind="port",
                sample=f"{n} Neel steps of the same 256x256 pinwheel workload in {dt:.1f} s, FP64 NumPy, "
                       f"1 thread (host has {cores} cores, {platform.processor() or platform.machine()})")


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every run here uses one MI355X).",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    import hotspice_amd as hs
    from hotspice_amd.workloads import neel_bytes_per_attempt, pinwheel_neel

    mm = pinwheel_neel(recompute_every=PERIOD)
    n = mm.n
    mc = PERIOD / n
    for _ in range(args.warmup):
        mm.progress(t_max=math.inf, MCsteps_max=mc)

    # error of the incremental FP32 + Kahan field vs an FP64 recompute, at the end of one period
    mm._device_recompute()
    mm.progress(t_max=math.inf, MCsteps_max=(PERIOD - 1) / n)  # PERIOD-1 steps: no recompute yet
    E_inc = mm.E
    mm._device_recompute()
    E_ref = mm.E
    rel_err = float(np.max(np.abs(E_inc - E_ref)) / np.max(np.abs(E_ref)))

    # timed region: K benchmark steps of PERIOD attempts + FP64 recompute each; HIP events on the
    # launch stream bracket every launch of the persistent Neel kernel (k_neel_team) inside it
    dev = mm._dev
    dev.launch_events = []
    a0, s0 = mm.attempted_switches, mm.switches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        mm.progress(t_max=math.inf, MCsteps_max=mc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts = mm.attempted_switches - a0
    switches = mm.switches - s0
    value = attempts / wall
    launches = dev.launch_events
    dev.launch_events = None
    launch_us = [e0.elapsed_time(e1) * 1e3 for e0, e1, _ in launches]
    neel_us = sum(launch_us) / len(launch_us)
    steps_per_launch = attempts / len(launches)

    bytes_per = neel_bytes_per_attempt(n, n, perp=True)
    bytes_launch = bytes_per * steps_per_launch
    achieved = bytes_launch / (neel_us * 1e-6) / 1e9
    line = {
        "metric": "switch attempts/s (plain vs batched/served) at 1e-5 rel E err; p50/p99 job lat",
        "value": value,
        "unit": "switch attempts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+kahan (fields), f64 (recompute, time)",
        "data": "synthetic (seeded random initial state and E_B)",
        "config": {"workload": "plain Neel: IP_Pinwheel 256x256 (32768 magnets), full 511x511 dipolar kernel "
                               "+ 1 mT Zeeman at 22.5 deg, T=300 K, E_B lognormal 60 kT/5%, FP64 recompute every "
                               "1024 switches", "attempts_per_step": PERIOD, "attempts": attempts,
                   "switches": switches, "rel_E_err_vs_fp64": rel_err, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_neel_team", "launches": len(launches), "avg_launch_us": neel_us,
                     "attempts_per_launch": steps_per_launch, "bytes_per_attempt": bytes_per,
                     "bytes_per_launch": bytes_launch},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
