"""Pins of the oracle that do not depend on the reference's own trajectories.

* Philox4x32-10 known-answer vectors (Salmon et al., SC'11 reference implementation).
* The reference's only unit test, tests/test_hexagon.py:11-25: one kagome hexagon in a vortex
  state has dipolar energy -1e-7 m^2/a^3 (29 + 20 sqrt(3)/9) on every magnet.
* Brute-force O(n^2) point-dipole pair sums (energies.py:233-241 pair formulas) on tiny grids.
* Closed forms of E_barrier (core.py:641-656) and calc_r (core.py:731-733).
* Incremental == full recompute after random flip sequences (analysis_dipolarkernel.py:49-55).
"""
import math

import numpy as np
import pytest

from oracle import lattices, model, philox


def test_philox_kat():
    def h(w):
        return [int(x) for x in w]
    assert h(philox.philox4x32(0, 0, 0, 0, 0, 0)) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert h(philox.philox4x32(f, f, f, f, f, f)) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert h(philox.philox4x32(0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344, 0xA4093822, 0x299F31D0)) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_ranges_and_mean():
    u = philox.uniform53(7, 3, philox.TAG_NEEL, np.arange(200000))
    assert u.min() >= 0 and u.max() < 1
    v = philox.uniform53(7, 3, philox.TAG_NEEL, np.arange(200000), open_zero=True)
    assert v.min() > 0 and v.max() <= 1
    assert np.all(v - u == 2.0 ** -53)
    assert abs(u.mean() - 0.5) < 4 * (1 / math.sqrt(12 * 200000))
    assert philox.randint(0.0, -5, 5) == -5 and philox.randint(1 - 2 ** -53, -5, 5) == 4


def _uniform_pattern(g):
    """Magnets._get_m_uniform(angle=0) (core.py:138-144)."""
    ang = 1e-6
    if g.in_plane:
        return g.occupation * (2 * ((g.orientation[:, :, 0] * math.cos(ang) + g.orientation[:, :, 1] * math.sin(ang)) >= 0) - 1)
    return np.ones((g.ny, g.nx)) * ((math.floor(ang / math.pi - .5) % 2) * 2 - 1)


@pytest.mark.parametrize("nx", range(5, 11))
@pytest.mark.parametrize("ny", range(3, 11))
def test_kagome_hexagon_known_answer(nx, ny):
    l = 470e-9
    m = 1.1278401e-15
    a = l / math.tan(30 * math.pi / 180)
    g = lattices.make("IP_Kagome", a, nx=nx, ny=ny)
    dd = model.Dipolar()
    mm = model.OracleMagnets(g, T=300, moment=m, pattern=_uniform_pattern(g), energies=[dd], m_perp_factor=0)
    mm.occupation[3:, :] = 0
    mm.m[3:, :] = 0
    mm.occupation[:, 5:] = 0
    mm.m[:, 5:] = 0
    mm.m[(0, 0, 1), (1, 3, 4)] *= -1
    mm.update_energy()
    E = -1e-7 * m * m / a ** 3 * (29 + 20 * math.sqrt(3) / 9)
    np.testing.assert_allclose(dd.E[mm.m.astype(bool)], E, rtol=1e-9)


def _pair_energy(g, m, moment, cutoff=np.inf):
    """Brute force: E_i = 1e-7 mu_i m_i sum_j mu_j m_j k(r_ij) with point dipoles (I = 0)."""
    ys, xs = np.nonzero(g.occupation)
    E = np.zeros((g.ny, g.nx))
    for i in range(len(ys)):
        yi, xi = ys[i], xs[i]
        tot = 0.0
        for j in range(len(ys)):
            if i == j:
                continue
            yj, xj = ys[j], xs[j]
            rx, ry = g.x[xj] - g.x[xi], g.y[yj] - g.y[yi]
            r = math.hypot(rx, ry)
            if r > cutoff:
                continue
            ux, uy = rx / r, ry / r
            if g.in_plane:
                o1 = g.orientation[yi, xi]
                o2 = g.orientation[yj, xj]
                k = o1[0] * o2[0] * (3 * ux * ux - 1) + o1[1] * o2[1] * (3 * uy * uy - 1) + 3 * ux * uy * (o1[0] * o2[1] + o1[1] * o2[0])
                kern = -k / r ** 3
            else:
                kern = 1 / r ** 3
            tot += kern * m[yj, xj] * moment
        E[yi, xi] = 1e-7 * moment * m[yi, xi] * tot
    return E


@pytest.mark.parametrize("lat,n,cut", [("OOP_Square", 6, np.inf), ("IP_Pinwheel", 8, np.inf), ("IP_Square", 8, 2.1e-6),
                                       ("IP_Kagome", 12, np.inf)])
def test_dipolar_full_recompute_brute_force(lat, n, cut):
    g = lattices.make(lat, 1e-6, n=n)
    rng = np.random.default_rng(5)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    dd = model.Dipolar(cutoff_range=cut)
    mm = model.OracleMagnets(g, T=300, moment=2e-16, pattern=pat, energies=[dd])
    Eb = _pair_energy(g, mm.m, 2e-16, cut)
    np.testing.assert_allclose(dd.E, Eb, rtol=1e-10, atol=1e-12 * np.abs(Eb).max())


def test_kernel_symmetry():
    """K_n[d] of the site at the centre equals K_n'[-d] of the partner site (pair symmetry)."""
    g = lattices.make("IP_Kagome", 1e-6, n=16)
    mm = model.OracleMagnets(g, T=300, moment=2e-16, pattern=1, energies=[])
    K, Kps, Kpo, idx = model.build_kernels(mm)
    ny, nx = g.ny, g.nx
    tol = 1e-12 * np.abs(K).max()
    for y in range(g.uy):
        for x in range(g.ux):
            n = idx[y, x]
            if n < 0:
                continue
            assert K[n, ny - 1, nx - 1] == 0
            for dy in range(-3, 4):
                for dx in range(-3, 4):
                    y2, x2 = (y + dy) % g.uy, (x + dx) % g.ux
                    n2 = idx[y2, x2]
                    if n2 < 0:
                        continue
                    assert K[n, ny - 1 + dy, nx - 1 + dx] == pytest.approx(K[n2, ny - 1 - dy, nx - 1 - dx], rel=1e-12, abs=tol)
                    # perpendicular partner: K_perpself of n at d == K_perpother of n2 at -d
                    assert Kps[n, ny - 1 + dy, nx - 1 + dx] == pytest.approx(Kpo[n2, ny - 1 - dy, nx - 1 - dx], rel=1e-12, abs=tol)


def test_barrier_closed_forms():
    g = lattices.make("OOP_Square", 1e-6, n=4)
    kT = model.kB * 300
    rng = np.random.default_rng(0)
    E_B = 20 * kT * (1 + 0.1 * rng.random((4, 4)))
    mm = model.OracleMagnets(g, T=300, E_B=E_B, moment=1e-16, pattern=1, energies=[model.Zeeman(magnitude=0.0)])
    z = mm.get_energy("zeeman")
    z.E = rng.normal(0, 30 * kT, (4, 4))
    dE = -2 * z.E
    b = mm.E_barrier(min_only=True)
    np.testing.assert_allclose(b, np.maximum(dE, E_B - z.E))
    mm.params["ENERGY_BARRIER_METHOD"] = "parabolic"
    b = mm.E_barrier(min_only=True)
    ref = np.where(dE > 4 * E_B, dE, np.minimum(E_B * (dE / (4 * E_B) + 1) ** 2, dE + 4 * E_B))
    np.testing.assert_allclose(b, ref)
    # at dE = 0 both formulas give E_B (symmetric double well)
    z.E[:] = 0
    np.testing.assert_allclose(mm.E_barrier(min_only=True), E_B)


def test_calc_r_closed_form():
    g = lattices.make("OOP_Square", 200e-9, n=16)
    mm = model.OracleMagnets(g, T=300, moment=1.1e-15, pattern=1, energies=[])
    r = (8e-7 * 1.1e-15 ** 2 / (0.05 * model.kB * 300)) ** (1 / 3)
    assert mm.calc_r(0.05) == pytest.approx(min(r, 2 * 15 * 200e-9), rel=1e-14)


@pytest.mark.parametrize("lat", ["OOP_Square", "IP_Pinwheel"])
def test_incremental_equals_full_recompute(lat):
    g = lattices.make(lat, 1e-6, n=16)
    rng = np.random.default_rng(3)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    kT = model.kB * 300
    ens = [model.Dipolar(), model.Zeeman(magnitude=1e-3, angle=.4), model.Exchange(J=kT)]
    mm = model.OracleMagnets(g, T=300, moment=2e-16, pattern=pat, energies=ens,
                             params=dict(REDUCED_KERNEL_SIZE=0, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3))
    occ = np.argwhere(g.occupation)
    for it in range(30):
        k = 1 if it % 3 == 0 else int(rng.integers(2, 12))
        sel = occ[rng.choice(len(occ), k, replace=False)].T
        mm.m[sel[0], sel[1]] *= -1
        mm.update_energy(sel)
    E_inc, Ep_inc = mm.E.copy(), mm.E_perp.copy() if mm.USE_PERP_ENERGY else None
    mm.update_energy()
    scale = np.abs(mm.E).max()
    np.testing.assert_allclose(E_inc, mm.E, rtol=0, atol=1e-12 * scale)
    if Ep_inc is not None:
        np.testing.assert_allclose(Ep_inc, mm.E_perp, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("widen,consistent", [(1.0, False), (1 + 1e-9, True)])
def test_cutoff_at_a_lattice_distance(widen, consistent):
    """A cutoff_range equal to a lattice distance (10 pitches on IP_Square 22x22): the FP64
    cumulative-sum offsets put +10 pitches a few ulp above the cutoff and -10 below it, so the
    kernel keeps only one of the two entries. The full recompute (correlation with K, as
    energies.py:303) and the incremental updates (K indexed by target - source, as
    energies.py:320 and 350) then differ by that entry; a cutoff widened by 1e-9 keeps both and
    the two agree again. This is the reason the cutoff workloads widen their cutoffs."""
    g = lattices.make("IP_Square", 1e-6, n=22)
    rng = np.random.default_rng(5)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    mm = model.OracleMagnets(g, T=300, moment=2.37e-16, pattern=pat,
                             energies=[model.Dipolar(cutoff_range=10 * 0.5e-6 * widen)],
                             params=dict(REDUCED_KERNEL_SIZE=10, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0))
    occ = np.argwhere(g.occupation)
    for it in range(10):
        sel = occ[rng.choice(len(occ), 6, replace=False)].T
        mm.m[sel[0], sel[1]] *= -1
        mm.update_energy(sel)
    E_inc = mm.E.copy()
    mm.update_energy()
    err = np.abs(E_inc - mm.E).max() / np.abs(mm.E).max()
    if consistent:
        assert err < 1e-12
    else:
        assert err > 1e-5


@pytest.mark.parametrize("lat,n,expect", [("IP_Pinwheel", 33, False), ("IP_Pinwheel", 32, True),
                                          ("OOP_Square", 33, True), ("IP_Square", 31, False)])
def test_pbc_refused_without_whole_unit_cells(lat, n, expect):
    """core.py:361-375 keeps PBC off unless nx and ny are multiples of the unit cell (2x2 for the
    square and pinwheel lattices, 1x1 for OOP_Square)."""
    om = model.make_magnets(lat, 1e-6, n=n, PBC=True, T=300, moment=1e-16,
                            pattern=1.0, energies=[model.Exchange(J=1e-21)])
    assert om.PBC is expect
