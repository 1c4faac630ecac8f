"""Lock-step parity of the HIP path with the oracle over seeded random configurations: lattice,
size, PBC, update scheme, barrier method, energy mix (dipolar with and without a cutoff, dumbbell
dipoles, Zeeman, exchange), temperature, Metropolis kernel size, sum-or-convolution cutoff and
sampling radius. The bars are those of the fixed cases (test_gpu_parity.lockstep)."""
import math

import numpy as np
import pytest

import cases
from test_gpu_parity import lockstep

pytestmark = pytest.mark.gpu

LATTICES = ["OOP_Square", "IP_Square", "IP_Pinwheel", "IP_Kagome", "IP_Ising"]


def _random_spec(rng, k):
    lat = str(rng.choice(LATTICES))
    n = int(rng.integers(20, 44)) if lat == "IP_Kagome" else int(rng.integers(10, 36))
    pbc = bool(rng.random() < 0.4) and lat != "IP_Kagome"
    scheme = "Néel" if rng.random() < 0.5 else "Metropolis"
    energies = []
    r = rng.random()
    if r < 0.45:
        cut = float(rng.choice([math.inf, 3.1e-6, 5.1e-6])) if not pbc else math.inf
        energies.append(("dipolar", dict(cutoff_range=cut) if math.isfinite(cut) else {}))
    elif r < 0.7 and lat.startswith("IP"):
        energies.append(("dimonopolar", dict(d=float(rng.uniform(150e-9, 350e-9)),
                                             small_d=float(rng.uniform(80e-9, 200e-9)))))
    if rng.random() < 0.6:
        energies.append(("zeeman", dict(magnitude=float(rng.uniform(0, 3e-3)), angle=float(rng.uniform(0, 6.28)))))
    if rng.random() < 0.4 or not energies:
        energies.append(("exchange", dict(J_kT=float(rng.uniform(0.05, 0.6)))))
    params = dict(UPDATE_SCHEME=scheme)
    if scheme == "Metropolis":
        params.update(REDUCED_KERNEL_SIZE=int(rng.choice([0, 3, 5, 20])),
                      SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=int(rng.choice([0, 2, 20])))
        if pbc and params["REDUCED_KERNEL_SIZE"] == 0:
            # With PBC and the full kernel, the reference's convolution branch counts a pair once
            # per periodic image it reaches; the sum branch is the one both sides share.
            params["SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF"] = 1000
        update_kw = dict(r=float(rng.uniform(1.5e-6, 6e-6)))
    else:
        if rng.random() < 0.3:
            params["ENERGY_BARRIER_METHOD"] = "parabolic"
        update_kw = dict(t_max=np.inf)
    moment = 1.1e-15 if lat == "OOP_Square" else 2.37e-16
    return dict(lattice=lat, a=1e-6, n=n, T=float(rng.uniform(200, 900)), moment=moment, energies=energies,
                params=params, update_kw=update_kw, nsteps=60, seed=1000 + k, kw=dict(PBC=pbc) if pbc else {})


def _random_presample_spec(rng, k):
    """Grids of 150-220 cells a side sampled at 1.0-1.5 lattice spacings: more samples per step
    than one selection workgroup has threads, so every step takes the presampled route (wide
    sampling launch, batched selection, tile marking). Windows stay small (REDUCED_KERNEL_SIZE
    2-6), so both update branches (sums below the switch cutoff, truncated convolutions above)
    are reached."""
    lat = str(rng.choice(LATTICES))
    n = 2 * int(rng.integers(75, 111))
    pbc = bool(rng.random() < 0.4) and lat != "IP_Kagome"
    a = 1e-6
    energies = []
    r = rng.random()
    if r < 0.5 or not lat.startswith("IP"):
        energies.append(("dipolar", dict(cutoff_range=float(rng.choice([2.1e-6, 3.1e-6, 4.1e-6])))))
    else:
        energies.append(("dimonopolar", dict(d=float(rng.uniform(150e-9, 350e-9)),
                                             small_d=float(rng.uniform(80e-9, 200e-9)))))
    if rng.random() < 0.6:
        energies.append(("zeeman", dict(magnitude=float(rng.uniform(0, 3e-3)), angle=float(rng.uniform(0, 6.28)))))
    if rng.random() < 0.5:
        energies.append(("exchange", dict(J_kT=float(rng.uniform(0.05, 0.6)))))
    params = dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=int(rng.choice([2, 3, 4, 6])),
                  SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=int(rng.choice([0, 20, 100000])))
    moment = 1.1e-15 if lat == "OOP_Square" else 2.37e-16
    return dict(lattice=lat, a=a, n=n, T=float(rng.uniform(200, 900)), moment=moment, energies=energies,
                params=params, update_kw=dict(r=float(rng.uniform(1.0, 1.5)) * a), nsteps=16, seed=2000 + k,
                kw=dict(PBC=pbc) if pbc else {})


def _random_metro_spec(rng, k):
    """Metropolis with the other selection options the oracle replays exactly: the 'single'
    scheme, METROPOLIS_TIME (step time from the samples' barriers), and a radius derived from Q
    (calc_r) instead of an explicit r."""
    sp = _random_spec(rng, 4000 + k)
    while sp["params"]["UPDATE_SCHEME"] != "Metropolis":
        sp = _random_spec(rng, 4000 + k)
    sp["params"]["MULTISAMPLING_SCHEME"] = "single" if rng.random() < 0.3 else "grid"
    sp["params"]["METROPOLIS_TIME"] = bool(rng.random() < 0.5)
    if rng.random() < 0.5:
        sp["update_kw"] = dict(Q=float(rng.uniform(0.01, 0.3)))
    return sp


N_CONFIGS = 48
N_PRESAMPLE = 12
N_METRO = 16
_rng = np.random.default_rng(20261203)
RANDOM = {f"random_{k:02d}": _random_spec(_rng, k) for k in range(N_CONFIGS)}
_rng = np.random.default_rng(20261204)
RANDOM.update({f"random_presample_{k:02d}": _random_presample_spec(_rng, k) for k in range(N_PRESAMPLE)})
_rng = np.random.default_rng(20261205)
RANDOM.update({f"random_metro_{k:02d}": _random_metro_spec(_rng, k) for k in range(N_METRO)})


@pytest.mark.parametrize("name", sorted(RANDOM))
def test_random_config_lockstep(name):
    """Every other configuration runs with a short FP64 recompute period (7-43 switches), so the
    periodic recompute lands inside the lock-step trajectory."""
    spec = dict(RANDOM[name])
    k = sorted(RANDOM).index(name)
    every = 7 + 9 * (k % 5) if k % 2 else None
    if every:
        spec["product_kw"] = dict(recompute_every=every)
    neel = spec["params"]["UPDATE_SCHEME"] == "Néel"

    def recomputed(mm, n):
        # update() counts switches (Neel) or steps (Metropolis) towards the period; a recompute
        # is due once the count passes it, unless truncated windows are part of the result
        counted = mm.switches if neel else spec["nsteps"]
        if every and mm._incremental_exact() and counted > every:
            assert n >= 1, f"{name}: no FP64 recompute after {counted} counted steps (period {every})"

    cases.SYNTH[name] = spec
    try:
        compared, ties = lockstep(name, check_every=8 if "presample" in name else 20, check=recomputed)
    finally:
        del cases.SYNTH[name]
    assert compared + len(ties) == RANDOM[name]["nsteps"]


def _random_batch(scheme, seed, count=6):
    rng = np.random.default_rng(seed)
    specs, k = {}, 0
    while len(specs) < count:
        sp = _random_spec(rng, 500 + k)
        k += 1
        if sp["params"]["UPDATE_SCHEME"] == scheme:
            specs[f"batch_{seed}_{len(specs)}"] = sp
    return specs


@pytest.mark.parametrize("scheme,seed", [("Néel", 1), ("Néel", 2), ("Metropolis", 3), ("Metropolis", 4)])
def test_random_batch_equals_alone(scheme, seed):
    """batch.progress_many over a random mix of instances (lattices, sizes, energies, PBC, in-plane
    next to out-of-plane) ends every instance bit-identical to stepping it alone with
    Magnets.progress: same spins, fields and compensations, time and counters."""
    from hotspice_amd import batch
    from test_gpu_team import _same, _snapshot
    specs = _random_batch(scheme, seed)
    cases.SYNTH.update(specs)
    try:
        kws = [sp["update_kw"] if scheme == "Metropolis" else {} for sp in specs.values()]
        alone = [cases.product_magnets(n, cases.load(n)) for n in specs]
        for m, kw in zip(alone, kws):
            m.progress(t_max=math.inf, MCsteps_max=1.5, **kw)
        mixed = [cases.product_magnets(n, cases.load(n)) for n in specs]
        res = batch.progress_many(mixed, t_max=math.inf, MCsteps_max=1.5, per_instance=kws)
        for m1, m2, r in zip(alone, mixed, res):
            _same(_snapshot(m1), _snapshot(m2))
            assert r[1] == pytest.approx(m1.MCsteps, rel=1e-12)
    finally:
        for n in specs:
            del cases.SYNTH[n]


def _relax_specs(seed, count):
    """Random configurations with a pair energy: without one, exchange between perpendicular
    neighbours (IP_Square, pinwheel) is a cancellation zero and the sign of a switch energy,
    which decides relax and minimize, would be rounding noise on both sides."""
    rng = np.random.default_rng(seed)
    out, k = {}, 0
    while len(out) < count:
        sp = _random_spec(rng, 700 + k)
        k += 1
        if any(kind in ("dipolar", "dimonopolar") for kind, _ in sp["energies"]):
            out[f"relax_{len(out):02d}"] = sp
    return out


RELAX_RANDOM = _relax_specs(31, 10)


@pytest.mark.parametrize("method", ["relax", "minimize"])
@pytest.mark.parametrize("name", sorted(RELAX_RANDOM))
def test_random_relax_matches_oracle(name, method):
    """Magnets.relax / minimize (core.py:763-845) on random configurations against the oracle
    with the same numpy.random draws: the same outcome (including the reference's IndexError on
    grids shorter than two unit cells), spins bit-exact, counters, fields at 1e-5 of their scale."""
    sp = RELAX_RANDOM[name]
    cases.SYNTH[name] = sp
    try:
        fx = cases.load(name)
        mm, om = cases.product_magnets(name, fx), cases.oracle_magnets(name, fx)
        outcome = []
        for m in (mm, om):
            np.random.seed(17)
            try:
                getattr(m, method)()
                outcome.append("ok")
            except IndexError:
                outcome.append("IndexError")
        assert outcome[0] == outcome[1]
        np.testing.assert_array_equal(mm.m, om.m)
        assert mm.switches == om.switches and mm.attempted_switches == om.attempted_switches
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * cases.energy_scale(om))
    finally:
        del cases.SYNTH[name]


def _wolff_spec(rng, k):
    lat = str(rng.choice(["OOP_Square", "IP_Ising", "IP_Kagome", "OOP_Square"]))
    n = int(rng.integers(20, 44)) if lat == "IP_Kagome" else int(rng.integers(8, 40))
    pbc = bool(rng.random() < 0.5) and lat != "IP_Kagome"
    energies = [("exchange", dict(J_kT=float(rng.uniform(0.1, 0.9))))]
    if rng.random() < 0.5:
        energies.append(("zeeman", dict(magnitude=float(rng.uniform(0, 3e-4)), angle=float(rng.uniform(0, 6.28)))))
    moment = 1.1e-15 if lat == "OOP_Square" else 2.37e-16
    return dict(lattice=lat, a=1e-6, n=n, T=float(rng.uniform(200, 900)), moment=moment, energies=energies,
                params=dict(UPDATE_SCHEME="Wolff"), update_kw={}, nsteps=30, seed=3000 + k,
                kw=dict(PBC=pbc) if pbc else {})


_rng = np.random.default_rng(41)
WOLFF_RANDOM = {f"wolff_{k:02d}": _wolff_spec(_rng, k) for k in range(8)}


@pytest.mark.parametrize("name", sorted(WOLFF_RANDOM))
def test_random_wolff_matches_oracle(name):
    """Wolff cluster steps (core.py:588-602, 711-718) on random exchange(+Zeeman) configurations:
    the same cluster every step, spins bit-exact, counters, fields at 1e-5 of their scale."""
    sp = WOLFF_RANDOM[name]
    cases.SYNTH[name] = sp
    try:
        fx = cases.load(name)
        mm, om = cases.product_magnets(name, fx), cases.oracle_magnets(name, fx)
        for s in range(sp["nsteps"]):
            np.testing.assert_array_equal(mm.update(), om.update(), err_msg=f"{name} step {s}")
        np.testing.assert_array_equal(mm.m, om.m)
        assert mm.switches == om.switches and mm.attempted_switches == om.attempted_switches
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * cases.energy_scale(om))
    finally:
        del cases.SYNTH[name]


def test_update_steps_at_zero_t_max():
    """Magnets.update takes its step with no stop test (core.py:679-709): with t_max = 0 each Neel
    step counts an attempt, adds t_max = 0 to t and switches nothing, on the device as in the
    oracle; a following update(t_max=inf) switches the same magnet on both."""
    name = "random_00"
    assert RANDOM[name]["params"]["UPDATE_SCHEME"] == "Néel"
    cases.SYNTH[name] = RANDOM[name]
    try:
        fx = cases.load(name)
        mm, om = cases.product_magnets(name, fx), cases.oracle_magnets(name, fx)
        for _ in range(3):
            assert mm.update(t_max=0.0) is None and om.update(t_max=0.0) is None
        assert mm.attempted_switches == om.attempted_switches == 3
        assert mm.switches == om.switches == 0 and mm.t == om.t == 0.0
        np.testing.assert_array_equal(np.asarray(mm.update(t_max=math.inf)), np.asarray(om.update(t_max=math.inf)))
        np.testing.assert_array_equal(mm.m, om.m)
    finally:
        del cases.SYNTH[name]
