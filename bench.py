"""Benchmark of the Hotspice stepping path on one MI355X.

Headline workload (`--workload neel`, the plain Neel configuration): in-plane pinwheel ASI
256x256 (32768 magnets), full 511x511 dipolar kernels + 1 mT Zeeman field, Neel-Arrhenius,
FP32 + Kahan fields with an FP64 recompute every 1024 switches. One benchmark step = 1024 Neel
switch attempts + the FP64 recompute, i.e. one recompute period, and the K timed steps are one
progress() call (1.02e6 switches at the default K = 1000); `value` = switch attempts/s.

Other workloads (one JSON line each, same keys):
  metro64     plain Metropolis: OOP square 64x64, full 127x127 kernel, grid r = 8a, float64 fields;
              step = 1 MC sweep
  metro2048   plain Metropolis: OOP square 2048x2048, 41x41 window + exchange, grid r = 20a;
              step = 1 MC sweep
  batched     64 IP_Square instances (n 32..480, own T and field) stepped together by Metropolis;
              step = 1 MC sweep of every instance; also times the same instances one by one
              through the plain path (`vs_plain`)
  neel_batched 256 pinwheel 32x32 instances, one persistent Neel launch; step = 1024 attempts
              of every instance; `vs_plain` against the same instances stepped one by one
  served      continuous batching of clocking jobs (Poisson arrivals, Kagome / pinwheel,
              Neel / Metropolis); reports jobs/s and p50/p99 latency beside attempts/s

The default run (`--workload all`) measures the plain Neel line above as `value` and adds, timed
in the same run, a `secondary` object with the other configurations of the metric: metro2048
(plain, large), metro64 (plain, small, float64), batched (with `vs_plain`) and served, at the
offered 200 jobs/s and at a rate below what the device sustains -- each with its own `roofline`
and `cpu_baseline`.

Prints ONE JSON line. `roofline` is for the dominant kernel: algorithmic bytes per launch /
average launch duration, measured with HIP events recorded on the launch stream around every
launch in the timed region. `cpu_baseline` times the FP64 NumPy restatement of the reference on
a bounded sample of the same workload: on one host core for the plain workloads, and on P
independent processes (one per physical core, at most 16) for the batched one.

    python bench.py [--workload W] [--steps K] [--warmup W] [--cpu-seconds S] [--no-cpu]
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
# HBM bytes per launch of k_neel_team on the 256x256 pinwheel workload (1024 attempts), a profile
# figure: two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) scaled by a 256 MiB copy of known
# size; the commands and counter values are in profiles/neel_team_pinwheel256_pmc.md
NEEL_TEAM_TRAFFIC_PER_LAUNCH = 94.8e6
NEEL_TEAM_TRAFFIC_SOURCE = "profile figure (rocprofv3 PMC FETCH_SIZE + WRITE_SIZE, profiles/neel_team_pinwheel256_pmc.md)"
# HBM bytes per device step of the 2048x2048 Metropolis workload (presample with the tile bitmap +
# update), a profile figure from the same two PMC passes: profiles/metro_pmc_oop2048_ipsq64.md
# HBM bytes per device step of the 2048x2048 workload's first two sweeps from its random state, from a
# rocprofv3 PMC profile; once the grid has settled, steps accept far fewer switches and move less
METRO2048_TRAFFIC_PER_STEP = 24.05e6
METRO2048_TRAFFIC_SOURCE = ("profile figure per device step of the first 2 sweeps from the random state (rocprofv3 "
                            "PMC, profiles/metro_pmc_oop2048_ipsq64.md); not applied to a run whose steps accept "
                            "fewer switches")
METRIC = "switch attempts/s (plain vs batched/served) at 1e-5 rel E err; p50/p99 job lat"
WORKLOADS = ("all", "neel", "metro64", "metro2048", "batched", "neel_batched", "served")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--workload", choices=WORKLOADS, default="all")
    p.add_argument("--steps", type=int, default=None)
    p.add_argument("--warmup", type=int, default=None)
    p.add_argument("--cpu-seconds", type=float, default=20.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-shadow-refresh", dest="shadow_refresh", action="store_false",
                   help="every periodic FP64 recompute is a full one (for comparison)")
    p.add_argument("--jobs", type=int, default=32, help="served: number of jobs")
    p.add_argument("--window", type=int, default=None,
                   help="served: dipolar cutoff in cell pitches with the matching REDUCED_KERNEL_SIZE (default: the "
                        "full kernels and default SimParams of the served configuration)")
    p.add_argument("--field-steps", type=int, default=100, help="served: field steps per job")
    p.add_argument("--rate", type=float, default=200.0, help="served: arrivals per second")
    p.add_argument("--no-overlap", dest="overlap", action="store_false",
                   help="served: the Neel and Metropolis launches of an iteration on one stream (default: two)")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                   help="first copy MB MiB between two device buffers (a known read + write volume that "
                        "scales the FETCH_SIZE / WRITE_SIZE counters of a profiling run)")
    return p.parse_args()


def cores_note():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return f"host has {cores} cores, {platform.processor() or platform.machine()}"


def base_line(args, value, steps, warmup, wall, dtype, config, unit="switch attempts/s"):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": wall / max(steps, 1) * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic (seeded random initial state and E_B)",
            "config": config}


def roofline(kernel, launches, bytes_total, extra=None):
    us = [e0.elapsed_time(e1) * 1e3 for e0, e1, *_ in launches]
    avg = sum(us) / len(us)
    per_launch = bytes_total / len(us)
    achieved = per_launch / (avg * 1e-6) / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": None, "kernel": kernel, "launches": len(us), "avg_launch_us": avg, "bytes_per_launch": per_launch}
    if extra:
        out.update(extra)
    return out


def rel_field_error(mm):
    """Largest |E_incremental - E_FP64| / max|E_FP64| after the steps since the last recompute."""
    import numpy as np
    E_inc = mm.E
    mm._device_recompute()
    E_ref = mm.E
    return float(np.max(np.abs(E_inc - E_ref)) / np.max(np.abs(E_ref)))


# ---- plain Neel (default) ---------------------------------------------------------------------

def cpu_baseline_neel(seconds):
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
    return dict(value=n / dt, unit="switch attempts/s", cores=1, kind="port",
                sample=f"{n} Neel steps of the same 256x256 pinwheel workload in {dt:.1f} s, FP64 NumPy, "
                       f"1 thread ({cores_note()})")


def bench_neel(args):
    import torch
    from hotspice_amd.workloads import neel_bytes_per_attempt, pinwheel_neel
    steps = 1000 if args.steps is None else args.steps
    warmup = 5 if args.warmup is None else args.warmup
    mm = pinwheel_neel(recompute_every=PERIOD)
    n = mm.n
    mc = PERIOD / n
    for _ in range(warmup):
        mm.progress(t_max=math.inf, MCsteps_max=mc)

    # error of the incremental FP32 + Kahan field vs an FP64 recompute, at the end of one period
    mm._device_recompute()
    mm.progress(t_max=math.inf, MCsteps_max=(PERIOD - 1) / n)  # PERIOD-1 steps: no recompute yet
    rel_err = rel_field_error(mm)

    # timed region: K benchmark steps of PERIOD attempts + FP64 recompute each; HIP events on the
    # launch stream bracket every launch of the persistent Neel kernel (k_neel_team) inside it
    dev = mm._dev
    dev.launch_events = []
    a0, s0 = mm.attempted_switches, mm.switches
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    # one progress call of K periods (1.02e6 switches at the default K = 1000): its launches and
    # recomputes are queued back to back, and the state is read when the call ends
    mm.progress(t_max=math.inf, MCsteps_max=steps * mc)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts = mm.attempted_switches - a0
    launches, dev.launch_events = dev.launch_events, None
    bytes_per = neel_bytes_per_attempt(n, n, perp=True)
    line = base_line(args, attempts / wall, steps, warmup, wall, "f32+kahan (fields), f64 (recompute, time)",
                     {"workload": "plain Neel: IP_Pinwheel 256x256 (32768 magnets), full 511x511 dipolar kernel "
                                  "+ 1 mT Zeeman at 22.5 deg, T=300 K, E_B lognormal 60 kT/5%, FP64 recompute every "
                                  f"1024 switches, one progress call of {steps} x 1024 switches",
                      "attempts_per_step": PERIOD, "attempts": attempts,
                      "switches": mm.switches - s0, "rel_E_err_vs_fp64": rel_err, "parallelism": "single GPU"})
    line["roofline"] = roofline("k_neel_team", launches, bytes_per * attempts,
                                {"attempts_per_launch": attempts / len(launches), "bytes_per_attempt": bytes_per})
    if abs(attempts / len(launches) - PERIOD) < 1:
        line["roofline"]["traffic"] = NEEL_TEAM_TRAFFIC_PER_LAUNCH
        line["roofline"]["traffic_source"] = NEEL_TEAM_TRAFFIC_SOURCE
    line["latency_roofline"] = neel_latency_roofline(mm, line["roofline"]["avg_launch_us"] / (attempts / len(launches)))
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_neel(args.cpu_seconds)
    return line


def neel_latency_roofline(mm, us_per_attempt):
    """The step's latency floor, measured live: one exchange round of the instance's team (G
    workgroups publish a candidate and poll each other's through device-scope atomics, nothing
    else), against the measured time per attempt of k_neel_team."""
    import torch
    from hotspice_amd import _lib, batch
    G = batch.neel_team_size(mm)
    torch.cuda.synchronize()
    floor, failed = _lib.exchange_round_us(G, rounds=20000)
    return {"bound": "latency", "floor_us": floor, "achieved_us": us_per_attempt, "frac": floor / us_per_attempt,
            "unit": "us per attempt", "team_workgroups": G, "failed_rounds": failed,
            "floor_what": f"one cross-workgroup exchange round of {G} workgroups (hs_exchange_round_us, 20000 rounds)"}


# ---- plain Metropolis -------------------------------------------------------------------------

def metro_window_cells(mm, idx_list, kw):
    """Mean number of occupied cells within the kernel window of at least one accepted switch,
    per step, over the given accepted-index arrays (2, a)."""
    import numpy as np
    occ = np.asarray(mm.occupation).astype(bool)
    tot = 0
    for idx in idx_list:
        mark = np.zeros(occ.shape, bool)
        for y, x in zip(*np.asarray(idx).reshape(2, -1)):
            mark[max(0, y - kw):y + kw + 1, max(0, x - kw):x + kw + 1] = True
        tot += int((mark & occ).sum())
    return tot / max(1, len(idx_list))


def cpu_baseline_metro(make, kw, seconds, label):
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import model
    p = make()
    om = model.make_magnets("OOP_Square", p["a"], n=p["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"],
                            pattern=p["pattern"], energies=p["energies"], params=p["params"], seed=1, rng="numpy")
    om.update(**kw)
    a0 = om.attempted_switches
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        om.update(**kw)
    dt = time.perf_counter() - t0
    n = om.attempted_switches - a0
    return dict(value=n / dt, unit="switch attempts/s", cores=1, kind="port",
                sample=f"{n} Metropolis attempts of {label} in {dt:.1f} s, FP64 NumPy, 1 thread ({cores_note()})")


def cpu_baseline_metro2048(seconds):
    """The 2048x2048 workload in the FP64 NumPy restatement, one core: full (4095^2) kernels as the
    reference builds them, and its own update_multiple branch rule, so a step with more than 20
    switches per sublattice convolves the whole grid's switch mask with the 41x41 window
    (energies.py:343-361) -- about 13 s per step on one core. At least one step, then steps until
    `seconds` have passed."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    from oracle import model
    from hotspice_amd.workloads import edge_cutoff, kB, oop_large_inputs
    p = oop_large_inputs()
    kT = kB * p["T"]
    t_build = time.perf_counter()
    om = model.make_magnets("OOP_Square", p["a"], n=p["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"],
                            pattern=p["pattern"], energies=[model.Dipolar(cutoff_range=edge_cutoff(20 * p["a"])),
                                                            model.Exchange(J=0.5 * kT)],
                            params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=20), seed=1, rng="numpy")
    t_build = time.perf_counter() - t_build
    a0 = om.attempted_switches
    t0 = time.perf_counter()
    steps = 0
    while steps == 0 or time.perf_counter() - t0 < seconds:
        om.update(r=20 * p["a"])
        steps += 1
    dt = time.perf_counter() - t0
    n = om.attempted_switches - a0
    return dict(value=n / dt, unit="switch attempts/s", cores=1, kind="port",
                sample=f"{steps} Metropolis steps ({n} attempts) of the same 2048x2048 workload in {dt:.1f} s, FP64 "
                       f"NumPy with full kernels (built in {t_build:.0f} s), 1 thread ({cores_note()})")


def bench_metro(args, large):
    import numpy as np
    import torch
    from hotspice_amd.workloads import oop_large, oop_metropolis, oop_metropolis_inputs, kB
    a = 200e-9
    if large:
        mm, kw, sweeps, warm = oop_large(), dict(r=20 * a), 200, 0.2
        mm.recompute_every = max(1, int(16 * mm.n / 2850))  # every ~16 sweeps (about 2850 samples per step)
        desc = ("plain Metropolis: OOP_Square 2048x2048 (4.19M magnets), a=200 nm, moment 1.1e-15, T=350 K, "
                "E_B N(30 kT, 5%), dipolar cutoff 20a (41x41 window, REDUCED_KERNEL_SIZE=20) + exchange J=0.5 kT, "
                "grid r=20a, FP32+Kahan, FP64 recompute every ~16 sweeps")
        kwin = 20
    else:
        mm, kw, sweeps, warm = oop_metropolis(fp64=True), dict(r=8 * a), 1000, 1
        desc = ("plain Metropolis: OOP_Square 64x64 (4096 magnets), a=200 nm, moment 1.1e-15, T=300 K, "
                "E_B N(20 kT, 5%), full 127x127 dipolar kernel, grid r=8a, float64 fields, increments and "
                "acceptance (Magnets(fp64_fields=True))")
        kwin = 10 ** 6
    steps = sweeps if args.steps is None else args.steps
    mm.progress(t_max=math.inf, MCsteps_max=warm, **kw)
    dev = mm._dev
    dev.launch_events = []
    a0, s0 = mm.attempted_switches, mm.switches
    step0 = mm._state().step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mm.progress(t_max=math.inf, MCsteps_max=float(steps), **kw)  # one run of `steps` MC sweeps
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts, accepted = mm.attempted_switches - a0, mm.switches - s0
    launches, dev.launch_events = dev.launch_events, None
    nsteps = mm._state().step - step0
    # window cells of the accepted switches (N_T), sampled over 20 single steps after the timed region
    sample = [np.asarray(mm.update(**kw)).reshape(2, -1) for _ in range(20)]
    acc_s = sum(i.shape[1] for i in sample)
    if large:
        nt_per_acc = metro_window_cells(mm, sample, kwin) * len(sample) / max(1, acc_s)
        bytes_total = attempts * 5 + accepted * nt_per_acc * 17
    else:  # full kernel: a step with any accepted switch touches every magnet
        frac = sum(i.shape[1] > 0 for i in sample) / len(sample)
        nt_per_acc = mm.n * frac * len(sample) / max(1, acc_s)
        bytes_total = attempts * 5 + nsteps * frac * mm.n * 17
    err = rel_field_error(mm)
    dtype = "f32+kahan (fields), f64 (acceptance)" if large else "f64 (fields, increments, acceptance)"
    line = base_line(args, attempts / wall, steps, 0, wall, dtype,
                     {"workload": desc, "attempts": attempts, "accepted": accepted, "device_steps": nsteps,
                      "rel_E_err_vs_fp64": err, "window_cells_per_accept": nt_per_acc, "parallelism": "single GPU"})
    kname = ("k_metro_presample + k_metro_update_one (per library call of the progress call)" if large
             else "k_metro_persist_one_f64 (per library call of the progress call)")
    line["roofline"] = roofline(kname, launches, bytes_total)
    if large:
        line["roofline"]["traffic"] = None
        line["roofline"]["traffic_first_sweeps_per_step"] = METRO2048_TRAFFIC_PER_STEP
        line["roofline"]["traffic_source"] = METRO2048_TRAFFIC_SOURCE
    if not args.no_cpu:
        if large:
            line["cpu_baseline"] = cpu_baseline_metro2048(args.cpu_seconds)
        else:
            def make():
                p = oop_metropolis_inputs()
                from oracle import model
                p.update(energies=[model.Dipolar(recompute="fft")],
                         params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0))
                return p
            line["cpu_baseline"] = cpu_baseline_metro(make, kw, args.cpu_seconds, "the same 64x64 workload")
    return line


# ---- batched ----------------------------------------------------------------------------------

def _cpu_batched_worker(job):
    """One CPU worker of the batched baseline: builds its share of the 64 IP_Square instances in the
    FP64 NumPy restatement and steps them round-robin (one Metropolis step each, Q = 0.05) for
    `seconds`. Returns (attempts, seconds)."""
    specs, seconds = job
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import model
    from hotspice_amd.workloads import ipsq_batch_inputs
    mags = []
    for sp in specs:
        p = ipsq_batch_inputs(sp)
        mags.append(model.make_magnets(
            "IP_Square", p["a"], n=p["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"], pattern=p["pattern"],
            energies=[model.Dipolar(cutoff_range=p["cutoff"]), model.Zeeman(*p["field"])],
            params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=p["reduced_k"],
                        SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0), seed=p["seed"], rng="numpy"))
    for m in mags:  # warm-up (calc_r cache, first convolutions)
        m.update()
    a0 = sum(m.attempted_switches for m in mags)
    t0 = time.perf_counter()
    i = 0
    while time.perf_counter() - t0 < seconds:
        mags[i % len(mags)].update()
        i += 1
    return sum(m.attempted_switches for m in mags) - a0, time.perf_counter() - t0


def cpu_baseline_batched(specs, seconds):
    """The batched workload on the host the way the reference parallelises it (ParallelJobs.py:64):
    P = physical cores (at most 16, the CPU share of one GPU) independent processes, each stepping
    its share of the instances with the FP64 NumPy restatement; aggregate attempts/s."""
    import multiprocessing as mp
    import psutil
    P = max(1, min(16, psutil.cpu_count(logical=False) or 1, len(specs)))
    with mp.get_context("spawn").Pool(P) as pool:
        res = pool.map(_cpu_batched_worker, [(specs[i::P], seconds) for i in range(P)])
    rate = sum(a / dt for a, dt in res)
    n = sum(a for a, _ in res)
    return dict(value=rate, unit="switch attempts/s", cores=P, kind="port",
                sample=f"{n} Metropolis attempts of the 64 instances (round-robin, one step at a time) in "
                       f"{seconds:.0f} s on {P} processes x 1 thread, FP64 NumPy ({cores_note()})")


def bench_batched(args):
    import numpy as np
    import torch
    from hotspice_amd import batch
    from hotspice_amd.workloads import ipsq_batch_instance, ipsq_batch_specs
    sweeps = 10 if args.steps is None else args.steps
    specs = ipsq_batch_specs()
    mags = [ipsq_batch_instance(s) for s in specs]
    batch.progress_many(mags, t_max=math.inf, MCsteps_max=0.5)
    dev = mags[0]._dev
    dev.launch_events = []
    a0 = sum(m.attempted_switches for m in mags)
    s0 = sum(m.switches for m in mags)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(sweeps):
        batch.progress_many(mags, t_max=math.inf, MCsteps_max=1.0)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts = sum(m.attempted_switches for m in mags) - a0
    accepted = sum(m.switches for m in mags) - s0
    launches, dev.launch_events = dev.launch_events, None
    # the same sweeps of a second copy of the batch (same seeds, so the same trajectories: the
    # acceptance rate, and with it the update work per attempt, falls as the spins relax) with the
    # step loop replayed from captured hipGraphs of 16 steps
    from hotspice_amd import _lib
    twin = [ipsq_batch_instance(s) for s in specs]
    batch.progress_many(twin, t_max=math.inf, MCsteps_max=0.5)
    g0 = sum(m.attempted_switches for m in twin)
    old = _lib.set_graph_steps(16)
    try:
        torch.cuda.synchronize()
        tg = time.perf_counter()
        for _ in range(sweeps):
            batch.progress_many(twin, t_max=math.inf, MCsteps_max=1.0)
        torch.cuda.synchronize()
        graph_rate = (sum(m.attempted_switches for m in twin) - g0) / (time.perf_counter() - tg)
    finally:
        _lib.set_graph_steps(old)
        # live executable graphs slow every later plain launch of the process (ROCm 7.2, MI355X:
        # the served mix measured after them ran at 60 % of its rate), so none outlives this line
        _lib.clear_graph_cache()
    twin_same = all(np.array_equal(a.m, b.m) for a, b in zip(mags[::8], twin[::8]))
    del twin
    err = max(rel_field_error(m) for m in mags[::8])
    # plain path: the same instances (fresh copies) stepped one at a time, one sweep each
    plain = [ipsq_batch_instance(s) for s in specs]
    for m in plain:
        m.progress(t_max=math.inf, MCsteps_max=0.5)
    torch.cuda.synchronize()
    p0 = sum(m.attempted_switches for m in plain)
    t1 = time.perf_counter()
    for m in plain:
        m.progress(t_max=math.inf, MCsteps_max=1.0)
    torch.cuda.synchronize()
    wall_plain = time.perf_counter() - t1
    plain_rate = (sum(m.attempted_switches for m in plain) - p0) / wall_plain
    # algorithmic bytes: k*(4+1) per attempt + 17 B per occupied window cell of each accepted
    # switch (21x21 windows, half of the IP_Square cells occupied)
    bytes_total = attempts * 5 + accepted * 21 * 21 * 0.5 * 17
    line = base_line(args, attempts / wall, sweeps, 1, wall, "f32+kahan (fields), f64 (acceptance)",
                     {"workload": "batched Metropolis: 64 IP_Square instances, n in {32..512} log-uniform (seed 3), "
                                  "T~U(100,400) K, B~U(0,5) mT at random angle, a=1 um, moment 2.37e-16, dipolar cutoff "
                                  "10 cells + REDUCED_KERNEL_SIZE=10, cutoff 0, grid multisampling Q=0.05, FP32+Kahan",
                      "instances": len(mags), "magnets": int(sum(m.n for m in mags)), "attempts": attempts,
                      "accepted": accepted, "rel_E_err_vs_fp64": err, "plain_attempts_per_s": plain_rate,
                      "vs_plain": attempts / wall / plain_rate, "graph_replay_attempts_per_s": graph_rate,
                      "graph_replay_same_spins": twin_same,
                      "parallelism": "single GPU"})
    line["roofline"] = roofline("k_metro_select_many + k_metro_update_many (per batched call)", launches, bytes_total)
    line["cpu_baseline"] = None if args.no_cpu else cpu_baseline_batched(specs, args.cpu_seconds)
    return line


def bench_neel_batched(args):
    import torch
    from hotspice_amd import batch
    from hotspice_amd.workloads import neel_bytes_per_attempt, pinwheel_neel
    steps = 20 if args.steps is None else args.steps
    count = 256
    mags = [pinwheel_neel(n=32, seed=100 + i, recompute_every=PERIOD) for i in range(count)]
    n = mags[0].n
    batch.progress_many(mags, t_max=math.inf, MCsteps_max=PERIOD / n)
    dev = mags[0]._dev
    dev.launch_events = []
    a0 = sum(m.attempted_switches for m in mags)
    # one progress call of `steps` recompute periods per instance (its launches and recomputes
    # are queued back to back; the states are read when the call ends)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.progress_many(mags, t_max=math.inf, MCsteps_max=steps * PERIOD / n)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    attempts = sum(m.attempted_switches for m in mags) - a0
    launches, dev.launch_events = dev.launch_events, None
    # error of the incremental fields at the end of a recompute period (PERIOD-1 switches after
    # the recompute each timed call ended with)
    batch.progress_many(mags[::32], t_max=math.inf, MCsteps_max=(PERIOD - 1) / n)
    err = max(rel_field_error(m) for m in mags[::32])
    one = mags[0]
    torch.cuda.synchronize()
    p0 = one.attempted_switches
    t1 = time.perf_counter()
    one.progress(t_max=math.inf, MCsteps_max=steps * PERIOD / n)  # the same call, alone
    torch.cuda.synchronize()
    plain_rate = (one.attempted_switches - p0) / (time.perf_counter() - t1)
    bytes_total = neel_bytes_per_attempt(n, n, perp=True) * attempts
    line = base_line(args, attempts / wall, steps, 1, wall, "f32+kahan (fields), f64 (recompute, time)",
                     {"workload": "batched Neel: 256 IP_Pinwheel 32x32 instances (512 magnets each), parameters of "
                                  "the 256x256 pinwheel workload, one progress call of `steps` x 1024 attempts per "
                                  "instance (a persistent launch per period), FP64 recompute every 1024 "
                                  "switches per instance", "instances": count, "attempts": attempts,
                      "rel_E_err_vs_fp64": err, "plain_attempts_per_s": plain_rate,
                      "vs_plain": attempts / wall / plain_rate, "parallelism": "single GPU"})
    line["roofline"] = roofline("k_neel_team (256 teams of 1 workgroup)", launches, bytes_total)
    return line


def _cpu_served_worker(specs, fields, works, counter, ready):
    """One CPU worker of the served baseline: the FP64 NumPy restatement of its jobs (at most two),
    their field steps played round-robin -- set_field, then 1000 update(t_max=inf) calls (Neel) or
    progress(t_max=inf, MCsteps_max=4) as its update() steps (Metropolis) -- until the parent
    stops it. After every step it adds the step's attempts to `counter`, so the parent reads the
    attempts completed within its window (a step cut off by the stop counts nothing)."""
    os.environ["OMP_NUM_THREADS"] = "1"
    from oracle import model
    from hotspice_amd.workloads import served_inputs
    mags = []
    for sp in specs:
        p = served_inputs(sp)
        dip = model.Dipolar(recompute="fft") if p["cutoff"] is None else model.Dipolar(cutoff_range=p["cutoff"])
        params = dict(UPDATE_SCHEME="Neel" if sp["scheme"] == "Néel" else "Metropolis")
        if p["reduced_k"] is not None:
            params["REDUCED_KERNEL_SIZE"] = p["reduced_k"]
        mags.append(model.make_magnets(sp["lattice"], p["a"], n=sp["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"],
                                       pattern=p["pattern"], energies=[dip, model.Zeeman(0.0, 0.0)], params=params,
                                       seed=sp["seed"], rng="numpy"))
    ready.release()
    k = 0
    while True:
        for m, sp, work in zip(mags, specs, works):
            mag, ang = fields[k % len(fields)]
            m.get_energy("zeeman").set_field(mag, ang)
            mc0 = m.MCsteps
            steps = 0
            while (steps < work) if sp["scheme"] == "Néel" else model.lower_than(m.MCsteps - mc0, work):
                a = m.attempted_switches
                m.update(t_max=math.inf) if sp["scheme"] == "Néel" else m.update()
                steps += 1
                with counter.get_lock():
                    counter.value += m.attempted_switches - a
        k += 1


def cpu_baseline_served(jobs, seconds):
    """The served jobs on the host the way the reference runs jobs (ParallelJobs.py:64): P = physical
    cores (at most 16) independent processes, each with at most two of the jobs. Once every worker
    has built its instances, the attempts completed in the next `seconds` are counted and the
    workers are stopped; aggregate attempts/s."""
    import multiprocessing as mp
    import psutil
    ctx = mp.get_context("spawn")
    P = max(1, min(16, psutil.cpu_count(logical=False) or 1, len(jobs)))
    shares = [jobs[i::P][:2] for i in range(P)]
    counter, ready = ctx.Value("q", 0), ctx.Semaphore(0)
    procs = [ctx.Process(target=_cpu_served_worker, daemon=True,
                         args=([j.spec for j in sh], sh[0].fields, [j.work for j in sh], counter, ready))
             for sh in shares]
    for p in procs:
        p.start()
    try:
        for _ in procs:
            if not ready.acquire(timeout=600):
                raise RuntimeError("a served CPU-baseline worker did not build its instances within 600 s")
        with counter.get_lock():
            c0 = counter.value
        t0 = time.perf_counter()
        time.sleep(seconds)
        with counter.get_lock():
            n = counter.value - c0
        dt = time.perf_counter() - t0
    finally:
        for p in procs:
            p.terminate()
        for p in procs:
            p.join(timeout=30)
    return dict(value=n / dt, unit="switch attempts/s", cores=P, kind="port",
                sample=f"{n} attempts completed in {dt:.0f} s by {P} processes x 1 thread, each stepping up to two of "
                       f"the jobs ({sum(len(sh) for sh in shares)} in all) through their field steps, FP64 NumPy "
                       f"({cores_note()})")


def serve_once(jobs, overlap=False):
    """Serves `jobs` on a fresh server; returns (done jobs, server, wall seconds, launch events)."""
    import torch
    from hotspice_amd._device import DeviceState
    from hotspice_amd.serve import Server
    srv = Server(overlap=overlap)
    for j in jobs:
        srv.submit(j)
    DeviceState.launch_events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    done = srv.run(timeout=900)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launches, DeviceState.launch_events = DeviceState.launch_events, None
    return done, srv, wall, launches


def served_roofline(done, launches):
    """Roofline of the served Neel launches (k_neel_team, every Neel job's team in them): the
    algorithmic bytes of the Neel jobs' attempts (full kernels: the update reads the whole grid)
    over the summed launch durations."""
    from hotspice_amd.workloads import _lattice_shape, neel_bytes_per_attempt
    neel = [l for l in launches if l[3] == "neel"]
    if not neel:
        return None
    total = 0.0
    for j in done:
        if j.spec["scheme"] != "Néel":
            continue
        ny, nx = _lattice_shape(j.spec["lattice"], j.spec["n"])
        n_occ = j.result["n"]
        w = j.spec.get("cutoff_cells")
        win = n_occ if w is None else min(n_occ, int((2 * w + 1) ** 2 * n_occ / (nx * ny)))
        total += j.result["attempts"] * neel_bytes_per_attempt(n_occ, win, perp=True)
    return roofline("k_neel_team (served Neel jobs)", neel, total)


def bench_served(args, rate=None, count=None, cpu=True):
    from hotspice_amd.serve import latency_stats
    from hotspice_amd.workloads import served_jobs
    rate = args.rate if rate is None else rate
    count = args.jobs if count is None else count
    jobs = served_jobs(count=count, rate=rate, field_steps=args.field_steps, cutoff_cells=args.window)
    done, srv, wall, launches = serve_once(jobs, overlap=args.overlap)
    st = latency_stats(done)
    win = "full kernels (default SimParams)" if args.window is None else \
        f"dipolar window {2 * args.window + 1}x{2 * args.window + 1} (cutoff {args.window} cells)"
    line = base_line(args, srv.attempts / wall, srv.iterations, 0, wall, "f32+kahan (fields), f64 (recompute, time)",
                     {"workload": f"served: {count} clocking jobs, Poisson arrivals at {rate:g}/s, 50% IP_Kagome "
                                  f"/ 50% IP_Pinwheel, n log-uniform 64..1024, 50% Neel (1000 attempts per field step) / "
                                  f"50% Metropolis (4 MC sweeps per field step), {args.field_steps} field steps "
                                  f"(1.5 mT (1-cos), 0/45 deg), {win}",
                      "jobs": len(done), "failed": len(srv.failed), "jobs_per_s": len(done) / wall,
                      "offered_jobs_per_s": rate, "latency_s": st, "attempts": srv.attempts,
                      "streams": 2 if args.overlap else 1, "parallelism": "single GPU"})
    line["roofline"] = served_roofline(done, launches)
    line["cpu_baseline"] = cpu_baseline_served(jobs, args.cpu_seconds / 2) if cpu and not args.no_cpu else None
    return line


def _secondary(line):
    """The part of a workload's line that the combined line carries for it."""
    keep = ("value", "unit", "ms_per_step", "steps", "roofline", "cpu_baseline")
    out = {k: line.get(k) for k in keep}
    out.update(line["config"])
    return out


def bench_all(args):
    """The plain Neel line (`value`), then the other configurations of the metric, each timed in
    this run with its own defaults: plain 2048x2048 Metropolis, the 64-instance batch with its
    plain-path comparison, 256 batched Neel instances with theirs, and the served mix at the
    offered 200 jobs/s and at half the rate the device sustained there."""
    t0 = time.perf_counter()

    def progress(what):
        print(f"bench: {what} done at {time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)

    line = bench_neel(args)
    progress("plain Neel")
    sub = argparse.Namespace(**vars(args))
    sub.steps = None
    sec = {}
    sec["metro2048"] = _secondary(bench_metro(sub, True))
    progress("plain Metropolis 2048x2048")
    small = argparse.Namespace(**vars(sub))
    small.cpu_seconds = args.cpu_seconds / 2
    sec["metro64"] = _secondary(bench_metro(small, False))
    progress("plain Metropolis 64x64 (float64)")
    sec["batched"] = _secondary(bench_batched(sub))
    progress("batched")
    sec["neel_batched"] = _secondary(bench_neel_batched(sub))
    progress("batched Neel")
    served = bench_served(sub, rate=args.rate, count=args.jobs)
    sec["served"] = _secondary(served)
    progress("served")
    sustain = served["config"]["jobs_per_s"]
    low = max(0.05, round(0.5 * sustain, 3))
    sec["served_below_saturation"] = _secondary(bench_served(sub, rate=low, count=max(8, args.jobs // 2), cpu=False))
    progress("served below saturation")
    sec["served_below_saturation"]["note"] = (f"offered {low:g} jobs/s, half of the {sustain:.3g} jobs/s the device "
                                              f"sustained at {args.rate:g}/s offered")
    line["secondary"] = sec
    return line


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every run here uses one MI355X).",
              file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        import torch
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fn = {"all": bench_all, "neel": bench_neel, "metro64": lambda a: bench_metro(a, False),
          "metro2048": lambda a: bench_metro(a, True), "batched": bench_batched, "neel_batched": bench_neel_batched,
          "served": bench_served}[args.workload]
    if not args.shadow_refresh:
        from hotspice_amd.core import Magnets
        Magnets.shadow_refresh = False
    line = fn(args)
    if not args.shadow_refresh:
        line["config"]["shadow_refresh"] = False
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
