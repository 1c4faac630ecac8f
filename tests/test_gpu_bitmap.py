"""Presampled Metropolis steps of one instance without PBC mark the tiles their accepted switches
reach in a bitmap (in the presampling launch), and the update launch reads it: no tile-listing
launch. Batched calls list the tiles with that launch instead, so the same instance stepped alone
and as a batch of one must end bit-identical (spins, compensated fields, t, counters), also when
the stop predicate ends a launch early and when the marked windows are wider than the truncated
ones the update applies."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

kB = 1.380649e-23


def _snapshot(mm):
    d = mm._dev
    out = [t.cpu().numpy().copy() for t in (d.m, d.E, d.Ec)]
    if mm.USE_PERP_ENERGY:
        out += [t.cpu().numpy().copy() for t in (d.Ep, d.Epc)]
    st = mm._state()
    return out + [st.t, st.attempts, st.switches, st.step]


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_array_equal(x, y)
        else:
            assert x == y


def _make(kind, seed):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    if kind == "oop_window_exchange":  # C5-like: 41x41 windows and exchange, about 2600 samples per step
        n, T = 512, 350
        kT = kB * T
        return hs.ASI.OOP_Square(200e-9, n, T=T, moment=1.1e-15, seed=seed, recompute_every=None,
                                 E_B=30 * kT * rng.normal(1, 0.05, (n, n)),
                                 pattern=np.where(rng.random((n, n)) < .5, -1., 1.),
                                 energies=[hs.DipolarEnergy(cutoff_range=20.5 * 200e-9), hs.ExchangeEnergy(J=0.5 * kT)],
                                 params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=20)), dict(r=5 * 200e-9)
    if kind == "oop_truncated":  # stored 41x41 windows, updates truncated to 11x11 above the cutoff
        n, T = 384, 500
        kT = kB * T
        return hs.ASI.OOP_Square(1e-6, n, T=T, moment=1.1e-15, seed=seed, recompute_every=None,
                                 E_B=20 * kT * rng.normal(1, 0.05, (n, n)),
                                 pattern=np.where(rng.random((n, n)) < .5, -1., 1.),
                                 energies=[hs.DipolarEnergy(cutoff_range=20.5e-6)],
                                 params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=5,
                                                     SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0)), dict(r=4e-6)
    # in-plane square lattice with E_perp (two sublattices, checkerboard occupancy)
    n, T = 192, 300
    probe = hs.ASI.IP_Square(1e-6, n, moment=2.37e-16, energies=[], params=hs.SimParams(UPDATE_SCHEME='Metropolis'))
    shape = (probe.ny, probe.nx)
    del probe
    return hs.ASI.IP_Square(1e-6, n, T=T, moment=2.37e-16, seed=seed, recompute_every=None,
                            E_B=20 * kB * T * rng.normal(1, 0.05, shape), pattern=np.where(rng.random(shape) < .5, -1., 1.),
                            energies=[hs.DipolarEnergy(cutoff_range=6.1e-6), hs.ZeemanEnergy(1e-3, 0.4)],
                            params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=6)), dict(r=1.5e-6)


@pytest.mark.parametrize("kind", ["oop_window_exchange", "oop_truncated", "ipsq_perp"])
def test_bitmap_steps_equal_listed_tiles(kind):
    from hotspice_amd import batch
    alone, kw = _make(kind, 5)
    twin, _ = _make(kind, 5)
    r = kw["r"]
    assert alone._select_geometry(r).nsc > 1024, "not a presampled step"
    for mc in (0.05, 0.13, 0.3):  # progress calls whose stop predicate ends their launches early
        d1 = alone.progress(t_max=math.inf, MCsteps_max=mc, **kw)
        d2 = batch.progress_many([twin], t_max=math.inf, MCsteps_max=mc, **kw)[0]
        assert d1[1] == d2[1] and d1[0] == d2[0]
        _same(_snapshot(alone), _snapshot(twin))
    assert alone.switches > 0
    if alone._incremental_exact():  # truncated updates are an approximation of the full sum by design
        E_inc = alone.E
        alone._device_recompute()
        assert float(np.max(np.abs(E_inc - alone.E)) / np.max(np.abs(alone.E))) < 1e-5
