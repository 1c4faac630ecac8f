"""Persistent team stepping: identical to the per-step launches, and an instance gives the same
result alone as inside a batch (bit-exact: same Philox streams, same arithmetic)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pinwheel(n, seed, **kw):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    EB = 60 * kT * np.exp(0.05 * rng.normal(size=(n, n)))
    return hs.ASI.IP_Pinwheel(1e-6, n, T=300, E_B=EB, moment=2.37e-16, pattern=pat, seed=seed,
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3, math.pi / 8)],
                              params=hs.SimParams(UPDATE_SCHEME='Néel'), **kw)


def _oop(n, seed, **kw):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    return hs.ASI.OOP_Square(1e-6, n, T=300, E_B=20 * kT, moment=1.1e-15, pattern=pat, seed=seed,
                             energies=[hs.DipolarEnergy(), hs.ExchangeEnergy(J=0.3 * kT)],
                             params=hs.SimParams(UPDATE_SCHEME='Néel', ENERGY_BARRIER_METHOD='parabolic'), **kw)


def _snapshot(mm):
    d = mm._dev
    return (d.m.cpu().numpy().copy(), d.E.cpu().numpy().copy(), d.Ec.cpu().numpy().copy(),
            d.Ep.cpu().numpy().copy(), mm.t, mm.attempted_switches, mm.switches)


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_array_equal(x, y)
        else:
            assert x == y


@pytest.mark.parametrize("make,n", [(_pinwheel, 32), (_oop, 24), (_oop, 23)])
@pytest.mark.parametrize("nblocks", [1, 2, 5, 0])
def test_team_equals_per_step_launches(make, n, nblocks):
    a = make(n, 3, recompute_every=None)
    b = make(n, 3, recompute_every=None)
    a.neel_persistent = False
    b.neel_blocks = nblocks
    a.progress(t_max=math.inf, MCsteps_max=1.5)
    b.progress(t_max=math.inf, MCsteps_max=1.5)
    _same(_snapshot(a), _snapshot(b))


def _oop_hot(n, seed, EB_kT, **kw):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    pat = np.where(rng.random((n, n)) < .5, -1., 1.)
    return hs.ASI.OOP_Square(1e-6, n, T=300, E_B=EB_kT * kT * rng.normal(1, 0.002, size=(n, n)), moment=1.1e-15,
                             pattern=pat, seed=seed, energies=[hs.DipolarEnergy(), hs.ExchangeEnergy(J=0.3 * kT)],
                             params=hs.SimParams(UPDATE_SCHEME='Néel'), **kw)


@pytest.mark.parametrize("EB_kT", [2000, 705])
@pytest.mark.parametrize("nblocks", [1, 3])
def test_team_extreme_barriers_equal_per_step(nblocks, EB_kT):
    """With t_max = +inf a team defers each step's waiting time unless the winning key says it
    could be huge, and then decides the time at once (its workgroup publishes it to the others).
    Barriers of ~2000 kT overflow every time to +inf (the first step ends a progress() call);
    ~705 kT give times near 1e290 that stay finite. Teams of one and of three workgroups equal
    the per-step launches bit for bit: spins, fields, t and counters."""
    a = _oop_hot(32, 5, EB_kT, recompute_every=None)
    b = _oop_hot(32, 5, EB_kT, recompute_every=None)
    a.neel_persistent = False
    b.neel_blocks = nblocks
    a.progress(t_max=math.inf, MCsteps_max=0.2)
    b.progress(t_max=math.inf, MCsteps_max=0.2)
    _same(_snapshot(a), _snapshot(b))
    if EB_kT == 2000:
        assert math.isinf(a.t) and a.attempted_switches == 1
    else:
        assert 1e250 < a.t < 1e305 and a.attempted_switches > 100
    for t_max in (math.inf, 1.0, math.inf, 1e-3, math.inf):
        ra, rb = a.update(t_max=t_max), b.update(t_max=t_max)
        assert (ra is None) == (rb is None)
        if ra is not None:
            assert tuple(map(int, ra)) == tuple(map(int, rb))
    _same(_snapshot(a), _snapshot(b))


def test_team_update_returns_index():
    a = _pinwheel(16, 4)
    b = _pinwheel(16, 4)
    b.neel_persistent = False
    for _ in range(30):
        ra, rb = a.update(t_max=math.inf), b.update(t_max=math.inf)
        assert tuple(map(int, ra)) == tuple(map(int, rb))
    _same(_snapshot(a), _snapshot(b))


def test_batch_equals_alone():
    from hotspice_amd import batch
    sizes = [(16, 11), (40, 12), (24, 13), (64, 14)]
    alone = [_pinwheel(n, s, recompute_every=64) for n, s in sizes]
    for m in alone:
        m.progress(t_max=math.inf, MCsteps_max=0.75)
    mixed = [_pinwheel(n, s, recompute_every=64) for n, s in sizes]
    res = batch.progress_many(mixed, t_max=math.inf, MCsteps_max=0.75, nblocks=[1, 3, 2, 0])
    for m1, m2, r in zip(alone, mixed, res):
        _same(_snapshot(m1), _snapshot(m2))
        assert r[1] == pytest.approx(m1.MCsteps, rel=1e-12)



METRO = ["metro_oop_window", "metro_ipsq_window", "metro_pinwheel", "metro_oop_conv"]


def _metro(name):
    from cases import load, product_magnets
    return product_magnets(name, load(name))


def test_metropolis_batch_equals_alone():
    """Every instance of a Metropolis batch (one selection launch with a workgroup per instance
    and one update launch over all instances' tiles per step) ends bit-identical to stepping it
    alone through Magnets.progress."""
    from cases import spec
    from hotspice_amd import batch
    kws = [spec(n)["update_kw"] for n in METRO]
    alone = [_metro(n) for n in METRO]
    for m, kw in zip(alone, kws):
        m.progress(t_max=math.inf, MCsteps_max=2.0, **kw)
    mixed = [_metro(n) for n in METRO]
    res = batch.progress_many(mixed, t_max=math.inf, MCsteps_max=2.0, per_instance=kws)
    for m1, m2, r in zip(alone, mixed, res):
        _same(_snapshot(m1), _snapshot(m2))
        assert r[1] == pytest.approx(m1.MCsteps, rel=1e-12)


def test_recompute_many_equals_single():
    """hs_recompute_fields_many gives every instance exactly the fields of its own
    hs_recompute_fields call."""
    mk = [lambda: _pinwheel(24, 5), lambda: _oop(20, 6), lambda: _pinwheel(40, 7)]
    a = [f() for f in mk]
    b = [f() for f in mk]
    for x, y in zip(a, b):
        x.progress(t_max=math.inf, MCsteps_max=0.3)
        y.progress(t_max=math.inf, MCsteps_max=0.3)
    for x in a:
        x._dev.recompute(3)
    b[0]._dev.recompute_many([y._dev for y in b], [3, 3, 1])
    for x, y, w in zip(a, b, [3, 3, 1]):
        np.testing.assert_array_equal(x._dev.E.cpu().numpy(), y._dev.E.cpu().numpy())
        np.testing.assert_array_equal(x._dev.Ec.cpu().numpy(), y._dev.Ec.cpu().numpy())
        if w == 3:
            np.testing.assert_array_equal(x._dev.Ep.cpu().numpy(), y._dev.Ep.cpu().numpy())
            np.testing.assert_array_equal(x._dev.Epc.cpu().numpy(), y._dev.Epc.cpu().numpy())


@pytest.fixture
def tune_cache(tmp_path, monkeypatch):
    from hotspice_amd import tune
    monkeypatch.setenv("HOTSPICE_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(tune, "_CACHE", {})
    monkeypatch.setattr(tune, "_LOADED", False)
    return tmp_path / "tune.json"


def test_autotuner_keeps_state_and_results(tune_cache):
    """tune_neel_blocks times team sizes on the instance itself, restores its state exactly and
    leaves a choice that gives the same trajectory as the default team."""
    from hotspice_amd import tune
    a = _pinwheel(64, 9)
    b = _pinwheel(64, 9)
    before = _snapshot(a)
    timing = tune.tune_neel_blocks(a, steps=64, reuse=False)
    assert set(timing) == set(tune.neel_candidates(a)) and a.neel_blocks in timing
    _same(_snapshot(a), before)
    a.progress(t_max=math.inf, MCsteps_max=0.5)
    b.progress(t_max=math.inf, MCsteps_max=0.5)
    _same(_snapshot(a), _snapshot(b))
    assert tune_cache.exists()
    c = _pinwheel(64, 10)
    c.neel_blocks = 'auto'  # a new instance of the same shape takes the cached choice
    c.progress(t_max=math.inf, MCsteps_max=0.1)
    assert c.neel_blocks == a.neel_blocks


def test_metropolis_autotuner_keeps_state_and_results(tune_cache):
    """tune_metro_blocks times update grids on the instance itself and restores its state; an
    instance with metro_blocks='auto' (alone and batched) steps bit-identically to the default
    grid, and a fixed grid does too."""
    from hotspice_amd import batch, tune
    from cases import spec
    kw = spec("metro_oop_presample")["update_kw"]
    a = _metro("metro_oop_presample")
    before = _snapshot(a)
    r = kw.get("r") or a.calc_r(kw.get("Q", 0.05))
    timing = tune.tune_metro_blocks(a, r, candidates=[1, 16, 175], steps=8, reuse=False)
    assert set(timing) == {1, 16, 175} and set(a._metro_tuned.values()) <= {1, 16, 175}
    _same(_snapshot(a), before)
    ref = _metro("metro_oop_presample")
    ref.progress(t_max=math.inf, MCsteps_max=0.5, **kw)
    a.metro_blocks = 'auto'
    a.progress(t_max=math.inf, MCsteps_max=0.5, **kw)
    _same(_snapshot(a), _snapshot(ref))
    fixed = _metro("metro_oop_presample")
    fixed.metro_blocks = 1
    fixed.progress(t_max=math.inf, MCsteps_max=0.5, **kw)
    _same(_snapshot(fixed), _snapshot(ref))
    many = [_metro("metro_oop_presample") for _ in range(2)]
    many[0].metro_blocks = 'auto'
    batch.progress_many(many, t_max=math.inf, MCsteps_max=0.5, **kw)
    for m in many:
        _same(_snapshot(m), _snapshot(ref))


def test_presampled_metropolis_batch_equals_alone():
    """A batch with instances of more than 1024 samples per step takes the presampled steps for
    all its instances (wide sampling and tile launches); every instance still ends bit-identical
    to stepping it alone."""
    from cases import spec
    from hotspice_amd import batch
    names = ["metro_oop_presample", "metro_ipsq_presample_pbc", "metro_oop_window"]
    kws = [spec(n)["update_kw"] for n in names]
    alone = [_metro(n) for n in names]
    for m, kw in zip(alone, kws):
        m.progress(t_max=math.inf, MCsteps_max=0.5, **kw)
    mixed = [_metro(n) for n in names]
    res = batch.progress_many(mixed, t_max=math.inf, MCsteps_max=0.5, per_instance=kws)
    for m1, m2, r in zip(alone, mixed, res):
        _same(_snapshot(m1), _snapshot(m2))
        assert r[1] == pytest.approx(m1.MCsteps, rel=1e-12)


@pytest.mark.parametrize("names", [["metro_oop_presample", "metro_ipsq_presample_pbc", "metro_oop_window"],
                                   ["metro_oop_presample", "metro_ipsq_pbc", "metro_pinwheel"]],
                         ids=["presampled", "listed"])
def test_metropolis_graph_replay_equals_plain_launches(names):
    """The batched Metropolis loop replays a captured hipGraph per group of 4 steps (two replays
    and two plain steps per call, the graph captured on the first call and reused after it, with
    the host reading the states and running recomputes between calls); after every call each
    instance is bit-identical to the same batch stepped with plain launches only. Halfway the
    graph cache is cleared (hs_graph_cache_clear), so the next call captures again."""
    from cases import spec
    from hotspice_amd import _lib, batch
    kws = [spec(n)["update_kw"] for n in names]

    def sels(mags):
        return [m._select_geometry(kw.get("r") or m.calc_r(kw.get("Q", 0.05))) for m, kw in zip(mags, kws)]

    plain = [_metro(n) for n in names]
    graphed = [_metro(n) for n in names]
    for m in plain + graphed:
        m._pre_step()
    sp, sg = sels(plain), sels(graphed)
    old = _lib.set_graph_steps(4)
    try:
        for call in range(4):
            if call == 2:
                _lib.clear_graph_cache()
            _lib.set_graph_steps(0)
            batch.launch_metropolis([(m, s, 10, math.inf, math.inf) for m, s in zip(plain, sp)], begin=call == 0)
            _lib.set_graph_steps(4)
            batch.launch_metropolis([(m, s, 10, math.inf, math.inf) for m, s in zip(graphed, sg)], begin=call == 0)
            for a, b in zip(plain, graphed):
                a._after_device_step()
                b._after_device_step()
                _same(_snapshot(a), _snapshot(b))
            assert all(m.attempted_switches > 0 for m in graphed)
    finally:
        _lib.set_graph_steps(old)
        _lib.clear_graph_cache()  # live graphs slow the plain launches of later tests


def test_varying_temperature_instance_in_a_neel_batch():
    """A Neel batch with an instance whose temperature varies over the grid: that instance steps
    through the per-step launches, the others in the team launch; each ends as it does alone."""
    from cases import load, product_magnets
    from hotspice_amd import batch

    def make():
        return [_pinwheel(32, 3), product_magnets("neel_pinwheel_tramp", load("neel_pinwheel_tramp")), _oop(24, 4)]

    alone = make()
    for m in alone:
        m.progress(t_max=math.inf, MCsteps_max=0.5)
    mixed = make()
    res = batch.progress_many(mixed, t_max=math.inf, MCsteps_max=0.5)
    for m1, m2, r in zip(alone, mixed, res):
        _same(_snapshot(m1), _snapshot(m2))
        assert r[1] == pytest.approx(m1.MCsteps, rel=1e-12)
