"""Pins of the oracle's zero-temperature relaxation (Magnets._minimize_all / relax,
hotspice/core.py:785-845).

* A sweep that switches nothing leaves every magnet at a local minimum: no magnet gains energy by
  switching (E <= 0 everywhere), and relax() stops there.
* Nearest-neighbour exchange only: the magnets of one subset (two unit cells apart) do not
  interact, so every sweep lowers E_tot by exactly 2 * sum of E of the magnets it switches; a
  spin switches only where more neighbours disagree with it than agree.
* The counters count every switch once.
"""
import numpy as np
import pytest

from oracle import lattices, model


def _random(g, seed):
    rng = np.random.default_rng(seed)
    return np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)


def test_exchange_subsets_lower_the_energy_exactly():
    g = lattices.make("OOP_Square", 1e-6, n=16)
    kT = model.kB * 300
    mm = model.OracleMagnets(g, T=300, moment=1e-16, pattern=_random(g, 1), energies=[model.Exchange(J=kT)])
    ex = mm._energies[0]
    np.random.seed(5)
    total_switches = 0
    for sweep in range(8):
        prev_m = mm.m.copy()
        # one subset at a time: its switches do not interact, so the drop is exactly 2 * sum(E)
        for i in np.random.permutation(4):
            y0, x0 = divmod(int(i), 2)
            m_before, E_before, tot_before = mm.m.copy(), ex.E.copy(), mm.E_tot
            mm._minimize_all(order=[i])
            sw = mm.m != m_before
            sub = np.zeros_like(sw)
            sub[y0::2, x0::2] = True
            assert not np.any(sw & ~sub)
            assert np.all(E_before[sw] > 0) and np.all(E_before[sub & ~sw] <= 0)
            assert mm.E_tot == pytest.approx(tot_before - 2 * E_before[sw].sum(), rel=1e-12, abs=1e-12 * kT)
            total_switches += int(sw.sum())
        assert mm.switches == mm.attempted_switches == total_switches
        if np.array_equal(prev_m, mm.m):
            break


@pytest.mark.parametrize("lat,n,pbc", [("IP_Square", 12, False), ("IP_Kagome", 32, False), ("OOP_Square", 12, True)])
def test_relax_ends_in_a_local_minimum(lat, n, pbc):
    g = lattices.make(lat, 1e-6, n=n, PBC=pbc)
    mm = model.OracleMagnets(g, T=300, moment=2e-16, pattern=_random(g, 3), energies=[model.Dipolar(),
                             model.Zeeman(2e-4, 0.3)], PBC=pbc)
    np.random.seed(11)
    s0 = mm.switches
    nsweeps = mm.relax()
    assert mm.switches == mm.attempted_switches and mm.switches > s0
    n_max = int(np.sqrt(g.nx ** 2 + g.ny ** 2))
    if nsweeps < n_max:  # stopped on its own: no magnet gains by switching, or a 2-cycle
        E = mm.E[g.occupation.astype(bool)]
        again = mm.switches
        mm._minimize_all()
        assert np.all(E <= 0) or mm.switches > again
        if np.all(E <= 0):
            assert mm.switches == again


@pytest.mark.parametrize("lat,n,ens", [("OOP_Square", 12, "exchange"), ("IP_Square", 10, "dipolar")])
def test_minimize_ends_in_a_local_minimum(lat, n, ens):
    g = lattices.make(lat, 1e-6, n=n)
    kT = model.kB * 300
    energies = [model.Exchange(J=kT), model.Zeeman(1e-5, 0.2)] if ens == "exchange" else [model.Dipolar()]
    mm = model.OracleMagnets(g, T=300, moment=2e-16, pattern=_random(g, 4), energies=energies)
    np.random.seed(3)
    tot0 = mm.E_tot
    mm.minimize()
    occ = g.occupation.astype(bool)
    assert mm.switches == mm.attempted_switches > 0
    if np.all(mm.E[occ] <= 0):  # ended because no magnet gains by switching
        s = mm.switches
        mm.minimize()
        assert mm.switches == s
    assert mm.E_tot < tot0

