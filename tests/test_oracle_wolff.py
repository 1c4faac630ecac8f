"""Pins of the oracle's Wolff cluster step (Magnets._update_Wolff / _select_cluster_exchange,
hotspice/core.py:588-602, 711-718).

* J/kT -> 0: no bond forms, the cluster is the seed alone.
* J/kT -> infinity: every bond forms, so the cluster is the connected region of spins aligned with
  the seed (nearest-neighbour connectivity, found here with scipy.ndimage.label, wrapped with PBC).
* Two aligned neighbours: the second joins with probability 1 - exp(-2 J/kT) (checked over many
  steps within 4.5 sigma of the binomial spread).
* Counters and spins: the whole cluster switches, and E_tot of an exchange-only system changes by
  the energy of the cluster boundary.
"""
import math

import numpy as np
import pytest
from scipy import ndimage

from oracle import lattices, model


def _wolff(n, J_kT, pattern, pbc=False, seed=1):
    g = lattices.make("OOP_Square", 1e-6, nx=n[1], ny=n[0], PBC=pbc)
    kT = model.kB * 300
    return model.OracleMagnets(g, T=300, moment=1e-16, pattern=pattern, energies=[model.Exchange(J=J_kT * kT)],
                               PBC=pbc, params=dict(UPDATE_SCHEME="Wolff"), seed=seed)


def test_no_bonds_at_weak_coupling():
    mm = _wolff((12, 12), 1e-12, 1)
    for _ in range(20):
        c = mm.update()
        assert c.shape == (2, 1)
    assert mm.switches == mm.attempted_switches == 20


@pytest.mark.parametrize("pbc", [False, True])
def test_strong_coupling_takes_the_aligned_domain(pbc):
    rng = np.random.default_rng(4)
    pat = np.where(rng.random((14, 14)) < .5, -1., 1.)
    mm = _wolff((14, 14), 1e6, pat, pbc=pbc)
    for _ in range(10):
        m0 = mm.m.copy()
        c = mm.update()
        seed = (c[0][0], c[1][0])
        cl = np.zeros(m0.shape, bool)
        cl[c[0], c[1]] = True
        # the component of the aligned spins that holds the cluster
        same = m0 == m0[seed]
        lab, _ = ndimage.label(same)
        if pbc:  # join labels across the wrapped edges
            for y in range(m0.shape[0]):
                if same[y, 0] and same[y, -1]:
                    lab[lab == lab[y, -1]] = lab[y, 0]
            for x in range(m0.shape[1]):
                if same[0, x] and same[-1, x]:
                    lab[lab == lab[-1, x]] = lab[0, x]
        comp = lab == lab[seed]
        np.testing.assert_array_equal(cl, comp)
        np.testing.assert_array_equal(mm.m, np.where(cl, -m0, m0))


def test_bond_probability():
    J_kT = 0.35
    p = 1 - math.exp(-2 * J_kT)
    mm = _wolff((1, 2), J_kT, 1)
    sizes = []
    for _ in range(3000):
        mm.m[:] = 1
        mm.update_energy()
        sizes.append(mm.update().shape[1])
    k = sum(s == 2 for s in sizes)
    assert set(sizes) <= {1, 2}
    assert abs(k / 3000 - p) < 4.5 * math.sqrt(p * (1 - p) / 3000)


def test_energy_change_is_the_cluster_boundary():
    rng = np.random.default_rng(7)
    pat = np.where(rng.random((12, 12)) < .5, -1., 1.)
    mm = _wolff((12, 12), 0.4, pat, pbc=True)
    ex = mm._energies[0]
    kT = model.kB * 300
    for _ in range(15):
        m0, tot0 = mm.m.copy(), mm.E_tot
        c = mm.update()
        cl = np.zeros(m0.shape, bool)
        cl[c[0], c[1]] = True
        # bonds between the cluster and the rest flip sign: dE = 2 J * sum over such bonds of m_i m_j
        d = 0.0
        for dy, dx in ((0, 1), (1, 0)):
            a, b = cl, np.roll(cl, (-dy, -dx), axis=(0, 1))
            mb = np.roll(m0, (-dy, -dx), axis=(0, 1))
            d += np.sum((a ^ b) * m0 * mb)
        assert mm.E_tot == pytest.approx(tot0 + 2 * ex.J * d, rel=1e-12, abs=1e-12 * kT)
