"""Synthetic workloads of the stepping benchmarks, as concrete inputs (seeded).

  pinwheel_neel   In-plane pinwheel ASI 256x256 (32768 magnets, checkerboard), a = 1 um,
                  moment 2.37e-16 A m^2, E_B lognormal (median 60 kT, sigma 5 %), T = 300 K,
                  Zeeman 1 mT at 22.5 deg, full 511x511 dipolar kernels (K, K_perpself,
                  K_perpother), Neel-Arrhenius with attempt time 1e-10 s, FP32 + Kahan fields,
                  FP64 recompute every 1024 switches.
  oop_metropolis  Out-of-plane square ASI 64x64 (4096 magnets), a = 200 nm, moment 1.1e-15,
                  E_B normal (mean 20 kT, sigma 5 %), T = 300 K, dipolar only with the full
                  127x127 kernel, Metropolis with grid multisampling r = 8a, random +-1 start.
  oop_large       Out-of-plane square ASI 2048x2048 (4.19M magnets), a = 200 nm, moment 1.1e-15,
                  E_B normal (30 kT, 5 %), T = 350 K, dipolar with cutoff 20a (41x41 window,
                  REDUCED_KERNEL_SIZE = 20; see edge_cutoff) + exchange J = 0.5 kT, Metropolis
                  grid r = 20a.
"""
import math

import numpy as np

kB = 1.380649e-23


def edge_cutoff(r):
    """A cutoff_range of r lattice-pitch multiples, widened by 1e-9 relative. The kernel's
    distances are FP64 cumulative sums of the cell pitches, so an offset of exactly r can land a
    few ulp above r on one side of the centre and below it on the other: the kernel then keeps
    the -r entry and drops the +r one. The reference's full recompute correlates with K (source -
    target) while its incremental updates use K (target - source), so such a kernel makes the two
    disagree by that entry. Widening the cutoff keeps both entries (the next lattice distance is
    far above 1e-9 relative)."""
    return r * (1 + 1e-9)


def pinwheel_neel_inputs(n=256, seed=0):
    rng = np.random.default_rng(seed)
    kT = kB * 300
    shape = (n, n)
    return dict(a=1e-6, n=n, T=300, moment=2.37e-16, seed=seed,
                E_B=60 * kT * np.exp(0.05 * rng.normal(0, 1, size=shape)),
                pattern=np.where(rng.random(shape) < 0.5, -1.0, 1.0),
                field=(1e-3, math.pi / 8))


def pinwheel_neel(n=256, seed=0, recompute_every=1024):
    import hotspice_amd as hs
    p = pinwheel_neel_inputs(n, seed)
    return hs.ASI.IP_Pinwheel(p["a"], n, T=p["T"], E_B=p["E_B"], moment=p["moment"], pattern=p["pattern"],
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(*p["field"])],
                              params=hs.SimParams(UPDATE_SCHEME='Néel'), seed=seed, recompute_every=recompute_every)


def oop_metropolis_inputs(n=64, seed=0):
    rng = np.random.default_rng(seed)
    kT = kB * 300
    shape = (n, n)
    return dict(a=200e-9, n=n, T=300, moment=1.1e-15, seed=seed,
                E_B=20 * kT * rng.normal(1, 0.05, size=shape),
                pattern=np.where(rng.random(shape) < 0.5, -1.0, 1.0))


def oop_metropolis(n=64, seed=0, fp64=False):
    """The 64x64 plain Metropolis configuration; fp64 holds and steps its field in float64, with
    no periodic recompute (as the reference's float64 arrays have none: there is no FP32 drift
    to reset)."""
    import hotspice_amd as hs
    p = oop_metropolis_inputs(n, seed)
    return hs.ASI.OOP_Square(p["a"], n, T=p["T"], E_B=p["E_B"], moment=p["moment"], pattern=p["pattern"],
                             energies=[hs.DipolarEnergy()], seed=seed, fp64_fields=fp64,
                             recompute_every=None if fp64 else 1024,
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis', MULTISAMPLING_SCHEME='grid',
                                                 REDUCED_KERNEL_SIZE=0))


def oop_large_inputs(n=2048, seed=0):
    rng = np.random.default_rng(seed)
    a, T = 200e-9, 350
    kT = kB * T
    shape = (n, n)
    return dict(a=a, n=n, T=T, moment=1.1e-15, seed=seed, E_B=30 * kT * rng.normal(1, 0.05, size=shape),
                pattern=np.where(rng.random(shape) < 0.5, -1.0, 1.0))


def oop_large(n=2048, seed=0):
    import hotspice_amd as hs
    p = oop_large_inputs(n, seed)
    kT = kB * p["T"]
    return hs.ASI.OOP_Square(p["a"], n, T=p["T"], E_B=p["E_B"], moment=p["moment"], pattern=p["pattern"], seed=seed,
                             energies=[hs.DipolarEnergy(cutoff_range=edge_cutoff(20 * p["a"])),
                                       hs.ExchangeEnergy(J=0.5 * kT)],
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=20))


def neel_bytes_per_attempt(n_occ, window_occ, perp=True):
    """Algorithmic HBM bytes of one Neel attempt: the barrier/argmin pass reads E, E_B (and
    E_perp) of every magnet (4 B each), the update pass read-modify-writes E and its Kahan
    compensation (8 + 8 B), reads m (1 B) and, with perp channels, E_perp + compensation (16 B)
    for every magnet in the kernel window."""
    return n_occ * (4 + 4 + (4 if perp else 0)) + window_occ * (8 + 8 + 1 + (16 if perp else 0))


# ---- batched and served workloads ------------------------------------------------------------------

def ipsq_batch_specs(count=64, seed=3):
    """64 independent in-plane square ASI: n_i in {32, 48, ..., 512} drawn log-uniform, T_i ~ U(100, 400) K,
    Zeeman B_i ~ U(0, 5) mT at theta_i ~ U(0, 2 pi); a = 1 um, moment 2.37e-16 A m^2 (borrowed from the
    pinwheel workload), dipolar cutoff 10 cell pitches with REDUCED_KERNEL_SIZE = 10 and
    SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF = 0, Metropolis grid with Q = 0.05."""
    rng = np.random.default_rng(seed)
    sizes = np.arange(32, 513, 16)
    logn = rng.uniform(np.log(32), np.log(512), size=count)
    ns = [int(sizes[np.argmin(np.abs(np.log(sizes) - v))]) for v in logn]
    out = []
    for i, n in enumerate(ns):
        out.append(dict(n=n, T=float(rng.uniform(100, 400)), B=float(rng.uniform(0, 5e-3)),
                        theta=float(rng.uniform(0, 2 * math.pi)), seed=int(seed * 1000 + i)))
    return out


def ipsq_batch_inputs(spec, a=1e-6):
    """The arrays and parameters of ipsq_batch_instance(spec) (same generator draws), for building
    the same instance elsewhere (the CPU baseline)."""
    n = spec["n"]
    rng = np.random.default_rng(spec["seed"])
    kT = kB * spec["T"]
    E_B = 20 * kT * rng.normal(1, 0.05, size=(n, n))
    pattern = np.where(rng.random((n, n)) < 0.5, -1.0, 1.0)
    return dict(a=a, n=n, T=spec["T"], E_B=E_B, moment=2.37e-16, pattern=pattern, cutoff=edge_cutoff(10 * a / 2),
                field=(spec["B"], spec["theta"]), reduced_k=10, seed=spec["seed"])


def ipsq_batch_instance(spec, a=1e-6):
    import hotspice_amd as hs
    n = spec["n"]
    rng = np.random.default_rng(spec["seed"])
    kT = kB * spec["T"]
    mm = hs.ASI.IP_Square(a, n, T=spec["T"], E_B=20 * kT * rng.normal(1, 0.05, size=(n, n)), moment=2.37e-16,
                          pattern=np.where(rng.random((n, n)) < 0.5, -1.0, 1.0), seed=spec["seed"],
                          energies=[hs.DipolarEnergy(cutoff_range=edge_cutoff(10 * a / 2)),
                                    hs.ZeemanEnergy(spec["B"], spec["theta"])],
                          params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=10,
                                              SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0))
    return mm


def clocking_fields(steps=100, peak=3e-3):
    """Field steps of the served jobs: magnitude 1.5 mT (1 - cos(2 pi i / steps)) (0 -> 3 -> 0 mT),
    angle 0 for even i and 45 deg for odd i."""
    return [(peak / 2 * (1 - math.cos(2 * math.pi * i / steps)), 0.0 if i % 2 == 0 else math.pi / 4)
            for i in range(steps)]


def served_jobs(count=32, rate=200.0, seed=4, n_range=(64, 1024), field_steps=100, neel_work=1000,
                metro_sweeps=4.0, cutoff_cells=None):
    """Jobs of the served workload: Poisson arrivals at `rate` jobs/s; 50 % IP_Kagome / 50 %
    IP_Pinwheel with n log-uniform in n_range; 50 % Neel (`neel_work` switch attempts per field
    step) / 50 % Metropolis (`metro_sweeps` MC sweeps per field step); `field_steps` clocking
    field steps; other parameters as the pinwheel workload (a = 1 um, moment 2.37e-16, T = 300 K,
    E_B lognormal median 60 kT / 5 %). Dipolar kernels are full, or cut off at `cutoff_cells` cell
    pitches with the matching REDUCED_KERNEL_SIZE."""
    from .serve import Job, footprint_bytes
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    fields = clocking_fields(field_steps)
    for i in range(count):
        t += float(rng.exponential(1.0 / rate))
        n = int(round(math.exp(rng.uniform(math.log(n_range[0]), math.log(n_range[1])))))
        lattice = "IP_Kagome" if rng.random() < 0.5 else "IP_Pinwheel"
        scheme = "Néel" if rng.random() < 0.5 else "Metropolis"
        spec = dict(lattice=lattice, n=n, scheme=scheme, seed=int(seed * 100000 + i), cutoff_cells=cutoff_cells)
        nx, ny = _lattice_shape(lattice, n)[::-1]
        n_sub = 6 if lattice == "IP_Kagome" else 2
        work = float(neel_work) if scheme == "Néel" else float(metro_sweeps)
        fp = footprint_bytes(nx, ny, n_sub) if cutoff_cells is None else \
            footprint_bytes(nx, ny, n_sub, kernel_full=False, kwx=cutoff_cells, kwy=cutoff_cells)
        jobs.append(Job(make=lambda spec=spec: served_instance(spec), fields=fields, work=work, arrival=t,
                        footprint=fp, name=f"{lattice} n={n} {scheme}", spec=spec))
    return jobs


def served_inputs(spec):
    """Arrays and parameters of one served job's instance (spec from served_jobs): E_B lognormal
    (median 60 kT, sigma 5 %) and a random +-1 start from the job's seed, a = 1 um, moment
    2.37e-16, T = 300 K, the dipolar cutoff (None = full kernels) in cell pitches."""
    r = np.random.default_rng(spec["seed"])
    kT = kB * 300
    shape = _lattice_shape(spec["lattice"], spec["n"])
    dx = 1e-6 / 4 if spec["lattice"] == "IP_Kagome" else 1e-6 / 2
    cut = spec.get("cutoff_cells")
    return dict(a=1e-6, T=300, moment=2.37e-16, E_B=60 * kT * np.exp(0.05 * r.normal(0, 1, size=shape)),
                pattern=np.where(r.random(shape) < 0.5, -1.0, 1.0),
                cutoff=None if cut is None else edge_cutoff(cut * dx), reduced_k=cut)


def served_instance(spec):
    """The Magnets of one served job (a Zeeman energy whose field the job's steps set)."""
    import hotspice_amd as hs
    p = served_inputs(spec)
    dip = hs.DipolarEnergy() if p["cutoff"] is None else hs.DipolarEnergy(cutoff_range=p["cutoff"])
    params = hs.SimParams(UPDATE_SCHEME=spec["scheme"]) if p["reduced_k"] is None else \
        hs.SimParams(UPDATE_SCHEME=spec["scheme"], REDUCED_KERNEL_SIZE=p["reduced_k"])
    return getattr(hs.ASI, spec["lattice"])(p["a"], spec["n"], T=p["T"], E_B=p["E_B"], moment=p["moment"],
                                            pattern=p["pattern"], seed=spec["seed"],
                                            energies=[dip, hs.ZeemanEnergy(0.0, 0.0)], params=params)


def _lattice_shape(lattice, n):
    """(ny, nx) of the lattices the served jobs use (ASI.py:371-382 for Kagome)."""
    if lattice == "IP_Kagome":
        return (int(n / math.sqrt(3)) // 4 * 4 - 1, n)
    return (n, n)
