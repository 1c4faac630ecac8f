"""Magnets(fp64_fields=True): the combined field is held in FP64 and every Metropolis step adds FP64
increments (K64 * (m mu64) sums and exchange terms), with the acceptance probability the FP64 exp of
the FP64 product, as the reference's float64 arrays do. Stepped in lock-step with the FP64 oracle,
the spins agree bit for bit at every step and the fields to FP64 rounding (1e-12 of the energy
scale, where the FP32 + Kahan path is held to 1e-5); the small plain configuration (64x64 OOP, full
127x127 kernel, grid r = 8a) is the first case. Other stepping paths refuse such an instance."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

kB = 1.380649e-23


def _inputs(n, seed, T=300, EB_kT=20):
    rng = np.random.default_rng(seed)
    kT = kB * T
    return dict(E_B=EB_kT * kT * rng.normal(1, 0.05, size=(n, n)), pattern=np.where(rng.random((n, n)) < .5, -1., 1.))


def _pair(kind, seed=3):
    """(device instance with FP64 fields, oracle instance, update() keywords)."""
    import hotspice_amd as hs
    from oracle import model
    if kind == "c1_shape":  # the small plain configuration's geometry and energies
        n, a, T, mom = 64, 200e-9, 300, 1.1e-15
        p = _inputs(n, seed, T)
        mm = hs.ASI.OOP_Square(a, n, T=T, moment=mom, seed=seed, fp64_fields=True, energies=[hs.DipolarEnergy()],
                               params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=0), **p)
        om = model.make_magnets("OOP_Square", a, n=n, T=T, moment=mom, seed=seed, energies=[model.Dipolar()],
                                params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0), **p)
        return mm, om, dict(r=8 * a)
    if kind == "oop_exchange_zeeman_window":  # truncated convolutions, exchange and a field
        n, a, T, mom = 40, 1e-6, 500, 1.1e-15
        p = _inputs(n, seed, T)
        kT = kB * T
        params = dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=3, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=2)
        mm = hs.ASI.OOP_Square(a, n, T=T, moment=mom, seed=seed, fp64_fields=True,
                               energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3), hs.ExchangeEnergy(J=0.3 * kT)],
                               params=hs.SimParams(**params), **p)
        om = model.make_magnets("OOP_Square", a, n=n, T=T, moment=mom, seed=seed,
                                energies=[model.Dipolar(), model.Zeeman(1e-3), model.Exchange(J=0.3 * kT)],
                                params=params, **p)
        return mm, om, dict(r=3e-6)
    # in-plane square lattice without E_perp, METROPOLIS_TIME, field at an angle
    n, a, T, mom = 20, 1e-6, 300, 2.37e-16
    probe = hs.ASI.IP_Square(a, n, moment=mom, energies=[], params=hs.SimParams(UPDATE_SCHEME='Metropolis'))
    shape = (probe.ny, probe.nx)
    del probe
    rng = np.random.default_rng(seed)
    p = dict(E_B=20 * kB * T * rng.normal(1, 0.05, size=shape), pattern=np.where(rng.random(shape) < .5, -1., 1.))
    params = dict(UPDATE_SCHEME="Metropolis", METROPOLIS_TIME=True)
    mm = hs.ASI.IP_Square(a, n, T=T, moment=mom, seed=seed, fp64_fields=True, m_perp_factor=0,
                          energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(2e-3, 0.7)], params=hs.SimParams(**params), **p)
    om = model.make_magnets("IP_Square", a, n=n, T=T, moment=mom, seed=seed, m_perp_factor=0,
                            energies=[model.Dipolar(), model.Zeeman(2e-3, 0.7)], params=params, **p)
    return mm, om, dict(r=2.5e-6)


def _scale(om):
    from cases import energy_scale
    return energy_scale(om)


@pytest.mark.parametrize("kind", ["c1_shape", "oop_exchange_zeeman_window", "ipsq_metropolis_time"])
def test_fp64_lockstep_with_oracle(kind):
    mm, om, kw = _pair(kind)
    scale = _scale(om)
    assert np.allclose(mm.E, om.E, rtol=0, atol=1e-12 * scale), "initial fields"
    switched = 0
    for s in range(200):
        a = np.asarray(mm.update(**kw)).reshape(2, -1)
        b = np.asarray(om.update(**kw)).reshape(2, -1)
        assert np.array_equal(a, b), f"{kind} step {s}: accepted switches differ"
        switched += a.shape[1]
        if s % 25 == 24:
            assert np.array_equal(mm.m, om.m)
            err = float(np.max(np.abs(mm.E - om.E))) / scale
            assert err < 1e-12, f"{kind} step {s}: field error {err:.3g} of the energy scale"
    assert switched > 0
    assert abs(mm.E_tot - om.E_tot) <= 1e-12 * abs(om.E_tot) + 1e-12 * scale
    if kind == "ipsq_metropolis_time":
        assert abs(mm.t - om.t) <= 1e-12 * abs(om.t)


def _c1_with(seed, **params):
    """The small configuration's geometry with other selection parameters."""
    import hotspice_amd as hs
    from oracle import model
    n, a, T, mom = 64, 200e-9, 300, 1.1e-15
    p = _inputs(n, seed, T)
    params = dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0, **params)
    mm = hs.ASI.OOP_Square(a, n, T=T, moment=mom, seed=seed, fp64_fields=True, energies=[hs.DipolarEnergy()],
                           params=hs.SimParams(**params), **p)
    om = model.make_magnets("OOP_Square", a, n=n, T=T, moment=mom, seed=seed, energies=[model.Dipolar()],
                            params=params, **p)
    return mm, om


def test_fp64_single_scheme_and_explicit_samples_match_oracle():
    """The 'single' scheme (core.py:489-494) and the caller's sample lists (update(idx=...),
    core.py:660-661) reach the FP64 acceptance through the selection's unloaded branch."""
    mm, om = _c1_with(8, MULTISAMPLING_SCHEME='single')
    scale = _scale(om)
    for s in range(300):
        a = np.asarray(mm.update()).reshape(2, -1)
        b = np.asarray(om.update()).reshape(2, -1)
        assert np.array_equal(a, b), f"single step {s}"
    assert mm.switches > 0 and np.array_equal(mm.m, om.m)
    assert float(np.max(np.abs(mm.E - om.E))) < 1e-12 * scale
    mm, om = _c1_with(9)
    rng = np.random.default_rng(1)
    for s, k in enumerate([1, 7, 300, 0, 64, 1000, 5]):
        flat = rng.choice(64 * 64, size=k, replace=False)
        idx = np.asarray([flat // 64, flat % 64])
        a = np.asarray(mm.update(idx=idx)).reshape(2, -1)
        b = np.asarray(om.update(idx=idx)).reshape(2, -1)
        assert np.array_equal(a, b), f"explicit samples, call {s}"
    assert np.array_equal(mm.m, om.m)
    assert float(np.max(np.abs(mm.E - om.E))) < 1e-12 * scale


def test_fp64_poisson_steps_keep_exact_fields():
    """Poisson-disc steps (no exact replay against the reference's sampler, a documented
    divergence) still add FP64 increments: the fields equal a fresh FP64 recompute."""
    mm, _ = _c1_with(10, MULTISAMPLING_SCHEME='Poisson')
    mm.recompute_every = None
    s0 = mm.switches
    for _ in range(200):
        mm.update(r=6 * 200e-9)
    assert mm.switches > s0
    E_inc = mm.E
    mm._device_recompute()
    assert float(np.max(np.abs(E_inc - mm.E)) / np.max(np.abs(mm.E))) < 1e-12


def test_fp64_progress_matches_oracle():
    """progress() with an MCsteps bound: same steps, spins and fields as the oracle's progress()."""
    mm, om, kw = _pair("c1_shape", seed=5)
    mm.recompute_every = None
    for _ in range(3):
        d1 = mm.progress(t_max=math.inf, MCsteps_max=2.0, **kw)
        d2 = om.progress(t_max=math.inf, MCsteps_max=2.0, **kw)
        assert math.isclose(d1[1], d2[1], rel_tol=1e-14)
        assert np.array_equal(mm.m, om.m)
    assert float(np.max(np.abs(mm.E - om.E))) < 1e-12 * _scale(om)


def test_fp64_fields_stay_exact_without_recompute():
    """Thousands of incremental FP64 steps with no recompute stay at FP64 rounding of a fresh
    FP64 recompute of the same spins (the FP32 + Kahan path drifts to ~1e-7 here)."""
    from hotspice_amd.workloads import oop_metropolis
    mm = oop_metropolis(fp64=True)
    mm.recompute_every = None
    mm.progress(t_max=math.inf, MCsteps_max=20.0, r=8 * 200e-9)
    assert mm.switches > 1000
    E_inc = mm.E
    mm._device_recompute()
    E_ref = mm.E
    assert float(np.max(np.abs(E_inc - E_ref)) / np.max(np.abs(E_ref))) < 1e-12


def test_fp64_exchange_only_matches_oracle():
    """An instance without a pair energy (exchange and Zeeman only) has no kernel arrays: its FP64
    steps add the exchange terms alone, in lock-step with the oracle."""
    import hotspice_amd as hs
    from oracle import model
    n, T = 32, 600
    kT = kB * T
    p = _inputs(n, 12, T)
    params = dict(UPDATE_SCHEME="Metropolis")
    mm = hs.ASI.OOP_Square(1e-6, n, T=T, moment=1.1e-15, seed=12, fp64_fields=True,
                           energies=[hs.ZeemanEnergy(8e-4), hs.ExchangeEnergy(J=0.2 * kT)],
                           params=hs.SimParams(**params), **p)
    om = model.make_magnets("OOP_Square", 1e-6, n=n, T=T, moment=1.1e-15, seed=12,
                            energies=[model.Zeeman(8e-4), model.Exchange(J=0.2 * kT)], params=params, **p)
    scale = _scale(om)
    for s in range(100):
        a = np.asarray(mm.update(r=4e-6)).reshape(2, -1)
        b = np.asarray(om.update(r=4e-6)).reshape(2, -1)
        assert np.array_equal(a, b), f"step {s}"
    assert mm.switches > 0 and np.array_equal(mm.m, om.m)
    assert float(np.max(np.abs(mm.E - om.E))) < 1e-12 * scale


def test_fp64_refused_where_not_supported():
    import hotspice_amd as hs
    from hotspice_amd import _lib
    p = _inputs(16, 1)
    neel = hs.ASI.OOP_Square(1e-6, 16, T=300, moment=1.1e-15, seed=1, fp64_fields=True, energies=[hs.DipolarEnergy()],
                             params=hs.SimParams(UPDATE_SCHEME='Néel'), **p)
    with pytest.raises(_lib.HotspiceError, match="FP32 fields only"):
        neel.update()
    p = _inputs(96, 2)
    big = hs.ASI.OOP_Square(1e-6, 96, T=300, moment=1.1e-15, seed=2, fp64_fields=True, energies=[hs.DipolarEnergy()],
                            params=hs.SimParams(UPDATE_SCHEME='Metropolis'), **p)
    with pytest.raises(_lib.HotspiceError, match="at most"):
        big.update(r=3e-6)
