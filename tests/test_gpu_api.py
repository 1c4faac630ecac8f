"""The drop-in API on the device: the reference's own unit test, field steps, progress(),
host edits of m, counters. Bars as in test_gpu_parity.py (FP32 + Kahan fields: 1e-5 of the
field scale; FP64 recomputes: 1e-9)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nx,ny", [(5, 3), (7, 4), (10, 10), (6, 9)])
def test_kagome_hexagon_known_answer(nx, ny):
    """tests/test_hexagon.py:11-25 of the reference, run through hotspice_amd."""
    import hotspice_amd as hotspice
    l = 470e-9
    m = 1.1278401e-15
    a = l / np.tan(30 * np.pi / 180)
    mm = hotspice.ASI.IP_Kagome(a, nx=nx, ny=ny, moment=m, energies=[dd := hotspice.energies.DipolarEnergy()], T=300,
                                m_perp_factor=0, PBC=False, seed=1)
    mm.occupation[3:, :] = mm.m[3:, :] = 0
    mm.occupation[:, 5:] = mm.m[:, 5:] = 0
    mm.m[(0, 0, 1), (1, 3, 4)] *= -1
    mm.update_energy()
    E = -1e-7 * m * m / a ** 3 * (29 + 20 * np.sqrt(3) / 9)
    DD_m = dd.E[mm.m.astype(bool)]
    assert np.allclose(DD_m, E, rtol=1e-9, atol=0)


def _pair(lat="IP_Kagome", n=20, seed=3, **kw):
    import hotspice_amd as hs
    from oracle import model
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    probe_geom = __import__("oracle.lattices", fromlist=["make"]).make(lat, 1e-6, n=n)
    shape = (probe_geom.ny, probe_geom.nx)
    pat = np.where(rng.random(shape) < .5, -1., 1.)
    EB = 40 * kT * rng.normal(1, 0.05, size=shape)
    mm = getattr(hs.ASI, lat)(1e-6, n, T=300, E_B=EB, moment=2.37e-16, pattern=pat, seed=seed,
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3, 0.3), hs.ExchangeEnergy(J=0.2 * kT)], **kw)
    om = model.make_magnets(lat, 1e-6, n=n, T=300, E_B=EB, moment=2.37e-16, pattern=pat, seed=seed,
                            energies=[model.Dipolar(), model.Zeeman(1e-3, 0.3), model.Exchange(J=0.2 * kT)],
                            params=dict(UPDATE_SCHEME="Neel"))
    return mm, om


def test_field_steps_match_oracle():
    mm, om = _pair()
    scale = np.abs(om.E).max()
    for mag, ang in [(2e-3, 0.0), (0.5e-3, math.pi / 4), (3e-3, 2.0), (0.0, 0.0)]:
        mm.get_energy('zeeman').set_field(mag, ang)
        om.get_energy('zeeman').set_field(mag, ang)
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(mm.E_perp, om.E_perp, rtol=0, atol=1e-6 * scale)
        np.testing.assert_allclose(mm.get_energy('zeeman').E, om.get_energy('zeeman').E, rtol=1e-12, atol=1e-30)
        assert mm.E_tot == pytest.approx(om.E_tot, rel=1e-6, abs=1e-6 * scale)


def test_per_energy_fields_match_oracle():
    mm, om = _pair()
    scale = np.abs(om.E).max()
    for name in ("dipolar", "zeeman", "exchange"):
        np.testing.assert_allclose(mm.get_energy(name).E, om.get_energy(name).E, rtol=0, atol=1e-9 * scale)
        np.testing.assert_allclose(mm.get_energy(name).E_perp, om.get_energy(name).E_perp, rtol=0, atol=1e-9 * scale)


def test_neel_progress_matches_oracle():
    mm, om = _pair(lat="IP_Pinwheel", n=16)
    dt1, mc1 = mm.progress(t_max=math.inf, MCsteps_max=1.0)
    dt2, mc2 = om.progress(t_max=math.inf, MCsteps_max=1.0)
    assert mc1 == mc2 and mm.attempted_switches == om.attempted_switches
    assert dt1 == pytest.approx(dt2, rel=1e-5)
    np.testing.assert_array_equal(mm.m, om.m)
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * scale)


def test_neel_progress_t_max():
    """Finite t_max: stops once the elapsed time reaches t_max (core.py:701-704, 756-757)."""
    mm, om = _pair(lat="IP_Pinwheel", n=16)
    dt1, mc1 = mm.progress(t_max=1e-6, MCsteps_max=math.inf)
    dt2, mc2 = om.progress(t_max=1e-6, MCsteps_max=math.inf)
    assert mm.attempted_switches == om.attempted_switches
    assert dt1 == pytest.approx(dt2, rel=1e-5)
    np.testing.assert_array_equal(mm.m, om.m)


def test_metropolis_progress_matches_oracle():
    import hotspice_amd as hs
    from oracle import model
    rng = np.random.default_rng(8)
    n = 40
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    kT = hs.kB * 300
    mm = hs.ASI.OOP_Square(1e-6, n, T=300, moment=1.1e-15, pattern=pat, seed=8,
                           energies=[hs.DipolarEnergy(cutoff_range=6e-6), hs.ExchangeEnergy(J=0.3 * kT)],
                           params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=6,
                                               SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3))
    om = model.make_magnets("OOP_Square", 1e-6, n=n, T=300, moment=1.1e-15, pattern=pat, seed=8,
                            energies=[model.Dipolar(cutoff_range=6e-6), model.Exchange(J=0.3 * kT)],
                            params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=6,
                                        SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3))
    r1 = mm.progress(MCsteps_max=3.0, r=3e-6)
    r2 = om.progress(MCsteps_max=3.0, r=3e-6)
    assert r1 == pytest.approx(r2)
    assert mm.attempted_switches == om.attempted_switches and mm.switches == om.switches
    np.testing.assert_array_equal(mm.m, om.m)
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * scale)


def test_host_edit_then_update_energy():
    mm, om = _pair()
    m = mm.m
    occ = np.argwhere(mm.occupation)[::7]
    m[occ[:, 0], occ[:, 1]] *= -1
    om.m[occ[:, 0], occ[:, 1]] *= -1
    mm.update_energy()
    om.update_energy()
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-9 * scale)
    np.testing.assert_array_equal(mm.m, om.m)


def test_counters_and_reset():
    mm, _ = _pair(lat="IP_Pinwheel", n=16)
    mm.progress(t_max=math.inf, MCsteps_max=0.25)
    assert mm.attempted_switches == math.ceil(0.25 * mm.n)
    assert 0 < mm.switches <= mm.attempted_switches and mm.t > 0
    mm.reset_stats()
    assert mm.t == 0 and mm.switches == 0 and mm.attempted_switches == 0 and mm.MCsteps == 0


def test_zero_temperature_returns_none():
    import hotspice_amd as hs
    mm = hs.ASI.OOP_Square(1e-6, 8, T=0, seed=1, params=hs.SimParams(UPDATE_SCHEME='Néel'))
    with pytest.warns(UserWarning):
        assert mm.update() is None


def test_pair_energy_swap_and_charges():
    """DiMonopolarEnergy replaces DipolarEnergy only after the latter is removed; the device pair
    weights follow (charges for the dumbbells, moments again afterwards), and the charges stay
    those taken at initialisation when the moment changes later (energies.py:413-415)."""
    import hotspice_amd as hs
    from oracle import model
    rng = np.random.default_rng(2)
    mm = hs.ASI.IP_Square(1e-6, 12, T=300, moment=2e-16, seed=4, energies=[hs.DipolarEnergy()])
    pat = mm.m.copy()
    with pytest.raises(NotImplementedError):
        mm.add_energy(hs.energies.DiMonopolarEnergy(d=300e-9))
    mm.remove_energy('dipolar')
    dm = mm.add_energy(hs.energies.DiMonopolarEnergy(d=300e-9, small_d=100e-9))
    assert mm.has_energy('dimonopolar') and not mm.has_energy('dipolar')
    om = model.make_magnets("IP_Square", 1e-6, n=12, T=300, moment=2e-16, pattern=pat,
                            energies=[model.DiMonopolar(d=300e-9, small_d=100e-9)])
    om.update_energy()
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-6 * scale)
    # a later moment change leaves the charges (and the dumbbell energies) as they were
    mm.moment = 2e-16 * (1 + rng.random(mm.m.shape))
    mm.update_energy()
    np.testing.assert_allclose(dm.E, om.E, rtol=0, atol=1e-6 * scale)
    # and back to point dipoles with the new moments
    mm.remove_energy('dimonopolar')
    mm.add_energy(hs.DipolarEnergy())
    od = model.make_magnets("IP_Square", 1e-6, n=12, T=300, moment=mm.moment, pattern=pat, energies=[model.Dipolar()])
    od.update_energy()
    np.testing.assert_allclose(mm.E, od.E, rtol=0, atol=1e-9 * np.abs(od.E).max())


RELAX_CASES = {
    "ipsq_dip_zee": dict(lat="IP_Square", n=16, kw={}, ens=[("dipolar", {}), ("zeeman", (3e-4, 0.4))], params={},
                         call={}),
    "kagome_mono_barrier": dict(lat="IP_Kagome", n=32, kw={}, ens=[("dimonopolar", dict(d=350e-9)), ("zeeman", (2e-3, 1.2))],
                                params={}, call=dict(ignore_barrier=False)),
    "oop_conv_window": dict(lat="OOP_Square", n=40, kw={}, ens=[("dipolar", {}), ("exchange", 0.5)],
                            params=dict(REDUCED_KERNEL_SIZE=5, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3), call={}),
    "ipsq_pbc_window": dict(lat="IP_Square", n=24, kw=dict(PBC=True), ens=[("dipolar", {}), ("zeeman", (5e-4, 0.2))],
                            params=dict(REDUCED_KERNEL_SIZE=4, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=2), call={}),
}


@pytest.mark.parametrize("name", sorted(RELAX_CASES))
def test_relax_matches_oracle(name):
    """Magnets.relax (core.py:808-845) on the device against the oracle, with the same subset
    orders (numpy.random seeded alike): spins bit-exact, fields at 1e-5 of their scale, counters."""
    import hotspice_amd as hs
    from oracle import lattices, model
    c = RELAX_CASES[name]
    kT = hs.kB * 300
    g = lattices.make(c["lat"], 1e-6, n=c["n"], PBC=c["kw"].get("PBC", False))
    rng = np.random.default_rng(9)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    EB = 2 * kT * rng.normal(1, 0.05, size=(g.ny, g.nx))
    pe, oe = [], []
    for kind, arg in c["ens"]:
        if kind == "dipolar":
            pe.append(hs.DipolarEnergy())
            oe.append(model.Dipolar())
        elif kind == "dimonopolar":
            pe.append(hs.energies.DiMonopolarEnergy(**arg))
            oe.append(model.DiMonopolar(**arg))
        elif kind == "zeeman":
            pe.append(hs.ZeemanEnergy(*arg))
            oe.append(model.Zeeman(*arg))
        else:
            pe.append(hs.ExchangeEnergy(J=arg * kT))
            oe.append(model.Exchange(J=arg * kT))
    mm = getattr(hs.ASI, c["lat"])(1e-6, c["n"], T=300, E_B=EB, moment=2.37e-16, pattern=pat, energies=pe, seed=1,
                                   params=hs.SimParams(**c["params"]), **c["kw"])
    om = model.make_magnets(c["lat"], 1e-6, n=c["n"], T=300, E_B=EB, moment=2.37e-16, pattern=pat, energies=oe,
                            params=dict(c["params"]), PBC=c["kw"].get("PBC", False))
    np.random.seed(21)
    mm.relax(**c["call"])
    np.random.seed(21)
    om.relax(**c["call"])
    assert mm.switches == om.switches > 0
    assert mm.attempted_switches == om.attempted_switches
    np.testing.assert_array_equal(mm.m, om.m)
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * scale)


def test_minimize_all_index_error_like_the_reference():
    """A subset origin outside a small grid raises IndexError after the subsets before it."""
    import hotspice_amd as hs
    mm = hs.ASI.IP_Kagome(1e-6, 16, T=300, moment=2.37e-16, seed=1, energies=[hs.DipolarEnergy()])
    assert mm.ny < 2 * mm.unitcell.y
    np.random.seed(0)
    with pytest.raises(IndexError):
        for _ in range(20):  # some permutation reaches an origin row >= ny
            mm._minimize_all()


@pytest.mark.parametrize("name", ["ipsq_dip_zee", "kagome_mono_barrier", "oop_conv_window"])
def test_minimize_matches_oracle(name):
    """Magnets.minimize (core.py:763-783) on the device against the oracle with the same
    numpy.random draws: spins bit-exact, counters, fields at 1e-5 of their scale."""
    import hotspice_amd as hs
    from oracle import lattices, model
    c = RELAX_CASES[name]
    kT = hs.kB * 300
    g = lattices.make(c["lat"], 1e-6, n=c["n"], PBC=c["kw"].get("PBC", False))
    rng = np.random.default_rng(19)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    EB = 2 * kT * rng.normal(1, 0.05, size=(g.ny, g.nx))
    pe, oe = [], []
    for kind, arg in c["ens"]:
        if kind == "dipolar":
            pe.append(hs.DipolarEnergy())
            oe.append(model.Dipolar())
        elif kind == "dimonopolar":
            pe.append(hs.energies.DiMonopolarEnergy(**arg))
            oe.append(model.DiMonopolar(**arg))
        elif kind == "zeeman":
            pe.append(hs.ZeemanEnergy(*arg))
            oe.append(model.Zeeman(*arg))
        else:
            pe.append(hs.ExchangeEnergy(J=arg * kT))
            oe.append(model.Exchange(J=arg * kT))
    mm = getattr(hs.ASI, c["lat"])(1e-6, c["n"], T=300, E_B=EB, moment=2.37e-16, pattern=pat, energies=pe, seed=1,
                                   params=hs.SimParams(**c["params"]), **c["kw"])
    om = model.make_magnets(c["lat"], 1e-6, n=c["n"], T=300, E_B=EB, moment=2.37e-16, pattern=pat, energies=oe,
                            params=dict(c["params"]), PBC=c["kw"].get("PBC", False))
    np.random.seed(5)
    mm.minimize(**c["call"])
    np.random.seed(5)
    om.minimize(**c["call"])
    assert mm.switches == om.switches > 0
    assert mm.attempted_switches == om.attempted_switches
    np.testing.assert_array_equal(mm.m, om.m)
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * scale)


def test_flip_cells_equals_update_energy():
    """hs_flip_cells (the switch + update_energy(index) of core.py:241-253) against the oracle's
    update_single / update_multiple for a single cell and a set of cells."""
    import hotspice_amd as hs
    from oracle import model
    kT = hs.kB * 300
    mm = hs.ASI.IP_Pinwheel(1e-6, 20, T=300, moment=2.37e-16, seed=3,
                            energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3, 0.5), hs.ExchangeEnergy(J=0.3 * kT)])
    om = model.make_magnets("IP_Pinwheel", 1e-6, n=20, T=300, moment=2.37e-16, pattern=mm.m,
                            energies=[model.Dipolar(), model.Zeeman(1e-3, 0.5), model.Exchange(J=0.3 * kT)])
    occ = np.argwhere(mm.occupation)
    rng = np.random.default_rng(2)
    for k in (1, 7, 40):
        sel = occ[rng.choice(len(occ), k, replace=False)]
        cells = sel[:, 0] * mm.nx + sel[:, 1]
        mm._dev.flip_cells(cells, mm.params.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF, mm.params.REDUCED_KERNEL_SIZE)
        mm._after_device_step()
        om.m[sel[:, 0], sel[:, 1]] *= -1
        om.update_energy(index=sel.T)
        np.testing.assert_array_equal(mm.m, om.m)
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * np.abs(om.E).max())
    with pytest.raises(IndexError):
        mm._dev.flip_cells([mm.nx * mm.ny], 20, 20)


@pytest.mark.parametrize("lat,n,pbc,extra", [("OOP_Square", 24, True, "zeeman"), ("OOP_Square", 40, False, None),
                                             ("IP_Square", 20, True, "dimonopolar")])
def test_wolff_matches_oracle(lat, n, pbc, extra):
    """Wolff cluster steps (core.py:588-602, 711-718) on the device against the oracle's Philox
    replay: the same clusters step by step, spins bit-exact, fields at 1e-5 of their scale."""
    import hotspice_amd as hs
    from oracle import lattices, model
    kT = hs.kB * 300
    g = lattices.make(lat, 1e-6, n=n, PBC=pbc)
    rng = np.random.default_rng(12)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    pe, oe = [hs.ExchangeEnergy(J=0.45 * kT)], [model.Exchange(J=0.45 * kT)]
    if extra == "zeeman":
        pe.append(hs.ZeemanEnergy(1e-4, 0.3))
        oe.append(model.Zeeman(1e-4, 0.3))
    elif extra == "dimonopolar":
        pe.append(hs.energies.DiMonopolarEnergy(d=300e-9))
        oe.append(model.DiMonopolar(d=300e-9))
    mm = getattr(hs.ASI, lat)(1e-6, n, T=300, moment=2.37e-16, pattern=pat, energies=pe, seed=6, PBC=pbc,
                              params=hs.SimParams(UPDATE_SCHEME='Wolff'))
    om = model.make_magnets(lat, 1e-6, n=n, T=300, moment=2.37e-16, pattern=pat, energies=oe, PBC=pbc, seed=6,
                            params=dict(UPDATE_SCHEME="Wolff"))
    sizes = []
    for s in range(40):
        a = mm.update()
        b = om.update()
        np.testing.assert_array_equal(a, b, err_msg=f"step {s}")
        sizes.append(a.shape[1])
    assert max(sizes) > 1
    np.testing.assert_array_equal(mm.m, om.m)
    assert mm.switches == om.switches and mm.attempted_switches == om.attempted_switches
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * np.abs(om.E).max())
    # progress() chunks the same steps on the device
    mm.progress(MCsteps_max=2.0)
    om.progress(MCsteps_max=2.0)
    np.testing.assert_array_equal(mm.m, om.m)
    assert mm.attempted_switches == om.attempted_switches


def test_wolff_needs_exchange_and_no_dipolar():
    import hotspice_amd as hs
    mm = hs.ASI.OOP_Square(1e-6, 8, T=300, seed=1, energies=[hs.DipolarEnergy()],
                           params=hs.SimParams(UPDATE_SCHEME='Wolff'))
    with pytest.raises(NotImplementedError):
        mm.update()
    mm = hs.ASI.OOP_Square(1e-6, 8, T=300, seed=1, energies=[hs.ZeemanEnergy(1e-3)],
                           params=hs.SimParams(UPDATE_SCHEME='Wolff'))
    with pytest.raises(NotImplementedError):
        mm.progress()


def test_update_after_time_overflow():
    """Barriers of 2000 kT overflow exp(beta * barrier) to +inf, so update(t_max=inf) switches a
    magnet and sets t = +inf. update(t_max=1) then compares the step's time with t_max itself
    (core.py:702-704): inf > 1, so t += 1 and nothing switches -- where t - t0 = inf - inf is NaN.
    Later update(t_max=inf) calls switch the same magnets as the oracle's."""
    import hotspice_amd as hs
    from oracle import model
    n = 12
    rng = np.random.default_rng(21)
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    kT = hs.kB * 300
    EB = 2000 * kT * rng.normal(1, 0.01, size=(n, n))
    mm = hs.ASI.OOP_Square(1e-6, n, T=300, E_B=EB, moment=1.1e-15, pattern=pat, seed=21,
                           energies=[hs.DipolarEnergy()], params=hs.SimParams(UPDATE_SCHEME='Néel'))
    om = model.make_magnets("OOP_Square", 1e-6, n=n, T=300, E_B=EB, moment=1.1e-15, pattern=pat, seed=21,
                            energies=[model.Dipolar()], params=dict(UPDATE_SCHEME="Neel"))
    for t_max in (math.inf, 1.0, 1.0, math.inf, 1e-3, math.inf, math.inf):
        a = mm.update(t_max=t_max)
        b = om.update(t_max=t_max)
        assert (a is None) == (b is None), (t_max, a, b)
        if b is not None:
            assert tuple(map(int, a)) == tuple(map(int, b))
        assert mm.t == om.t == math.inf
    assert mm.attempted_switches == om.attempted_switches == 7
    assert mm.switches == om.switches == 4
    np.testing.assert_array_equal(mm.m, om.m)


def test_metropolis_update_runs_the_periodic_recompute():
    """update() steps under Metropolis count towards the FP64 recompute period as progress() steps
    do: with recompute_every = 5, ten update() calls run two recomputes, and the fields stay at
    the FP64 oracle's."""
    import hotspice_amd as hs
    from oracle import model
    n = 24
    rng = np.random.default_rng(4)
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    # the full kernel in both update_multiple branches (REDUCED_KERNEL_SIZE = 0), so incremental
    # updates equal a recompute and the periodic one is due
    mm = hs.ASI.OOP_Square(1e-6, n, T=300, moment=1.1e-15, pattern=pat, seed=4, energies=[hs.DipolarEnergy()],
                           params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=0), recompute_every=5)
    om = model.make_magnets("OOP_Square", 1e-6, n=n, T=300, moment=1.1e-15, pattern=pat, seed=4,
                            energies=[model.Dipolar()], params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0))
    r0 = mm.recomputes
    for _ in range(10):
        np.testing.assert_array_equal(mm.update(r=4e-6), om.update(r=4e-6))
    assert mm.recomputes - r0 == 2
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-9 * np.abs(om.E).max())


@pytest.mark.parametrize("G", [1, 8, 64])
def test_exchange_round_probe(G):
    """hs_exchange_round_us: every round's records arrive, and a round costs less than a Neel
    step of a team of that size (it is that step's exchange alone)."""
    from hotspice_amd import _lib
    us, failed = _lib.exchange_round_us(G, rounds=2000)
    assert failed == 0
    assert 0.0 < us < 5.0
