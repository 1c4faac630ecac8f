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
                  REDUCED_KERNEL_SIZE = 20) + exchange J = 0.5 kT, Metropolis grid r = 20a.
"""
import math

import numpy as np

kB = 1.380649e-23


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


def oop_metropolis(n=64, seed=0):
    import hotspice_amd as hs
    p = oop_metropolis_inputs(n, seed)
    return hs.ASI.OOP_Square(p["a"], n, T=p["T"], E_B=p["E_B"], moment=p["moment"], pattern=p["pattern"],
                             energies=[hs.DipolarEnergy()], seed=seed,
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis', MULTISAMPLING_SCHEME='grid',
                                                 REDUCED_KERNEL_SIZE=0))


def oop_large(n=2048, seed=0):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    a, T = 200e-9, 350
    kT = kB * T
    shape = (n, n)
    return hs.ASI.OOP_Square(a, n, T=T, E_B=30 * kT * rng.normal(1, 0.05, size=shape), moment=1.1e-15,
                             pattern=np.where(rng.random(shape) < 0.5, -1.0, 1.0), seed=seed,
                             energies=[hs.DipolarEnergy(cutoff_range=20 * a), hs.ExchangeEnergy(J=0.5 * kT)],
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=20))


def neel_bytes_per_attempt(n_occ, window_occ, perp=True):
    """Algorithmic HBM bytes of one Neel attempt: the barrier/argmin pass reads E, E_B (and
    E_perp) of every magnet (4 B each), the update pass read-modify-writes E and its Kahan
    compensation (8 + 8 B), reads m (1 B) and, with perp channels, E_perp + compensation (16 B)
    for every magnet in the kernel window."""
    return n_occ * (4 + 4 + (4 if perp else 0)) + window_occ * (8 + 8 + 1 + (16 if perp else 0))
