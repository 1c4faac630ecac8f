"""Pins of the oracle's DiMonopolarEnergy (hotspice/energies.py:394-751).

* Closed forms of two dumbbells on a line and side by side (the plain kernel), and the zero
  interaction of a turned magnet on the axis of another (the perpendicular kernels).
* Brute force: every magnet an explicit pair of charges +-q = +-moment/d at its physical position;
  E_i = 1e-7 * prefactor * q_i^2 * m_i * sum_j m_j * (sum of the four charge-pair 1/r terms), with
  non-uniform moments and per-site d (the receiving magnet's charge only, energies.py:654-659),
  for open and periodic boundaries (one image per axis, as energies.py:610-623 folds the kernel).
* Far field: with d much smaller than the pitch the dumbbell kernels times q^2 approach the
  point-dipole kernels times moment^2.
* Incremental updates (update_single / update_multiple) equal the full recompute.
"""
import math

import numpy as np
import pytest

from oracle import lattices, model


def _charges(x, y, ang, d):
    h = d / 2
    return ((x + h * math.cos(ang), y + h * math.sin(ang), 1.0), (x - h * math.cos(ang), y - h * math.sin(ang), -1.0))


def _coulomb(A, B):
    return sum(qa * qb / math.hypot(xa - xb, ya - yb) for xa, ya, qa in A for xb, yb, qb in B)


def _coulomb_abs(A, B):
    return sum(1 / math.hypot(xa - xb, ya - yb) for xa, ya, qa in A for xb, yb, qb in B)


def _brute_energy(g, m, moment, dsite, small_d, pbc=False, pf=1.0):
    """E and E_perp of every magnet from explicit charges (see the module docstring), and the
    same sum over absolute values of the terms of E (the scale of its FP32 rounding)."""
    ys, xs = np.nonzero(g.occupation)
    Lx, Ly = float(np.sum(g.dx)), float(np.sum(g.dy))
    E = np.zeros((g.ny, g.nx))
    Ep = np.zeros((g.ny, g.nx))
    Eabs = np.zeros((g.ny, g.nx))

    def shifts(o):
        return (0,) if o == 0 else ((0, -1) if o > 0 else (0, 1))

    for i in range(len(ys)):
        yi, xi = ys[i], xs[i]
        ai, di = g.angles[yi, xi], dsite[yi, xi]
        Di = _charges(g.x[xi], g.y[yi], ai, di)
        Dp = _charges(g.x[xi], g.y[yi], ai + math.pi / 2, small_d)
        tot = totp = tabs = 0.0
        for j in range(len(ys)):
            if i == j:
                continue
            yj, xj = ys[j], xs[j]
            for sx in (shifts(xj - xi) if pbc else (0,)):
                for sy in (shifts(yj - yi) if pbc else (0,)):
                    Dj = _charges(g.x[xj] + sx * Lx, g.y[yj] + sy * Ly, g.angles[yj, xj], dsite[yj, xj])
                    tot += m[yj, xj] * _coulomb(Di, Dj)
                    totp += m[yj, xj] * _coulomb(Dp, Dj)
                    tabs += _coulomb_abs(Di, Dj)
        q2 = (moment[yi, xi] / di) ** 2
        E[yi, xi] = 1e-7 * pf * q2 * m[yi, xi] * tot
        Ep[yi, xi] = 1e-7 * pf * q2 * m[yi, xi] * totp
        Eabs[yi, xi] = 1e-7 * pf * q2 * tabs
    return E, Ep, Eabs


def _random_state(g, seed):
    rng = np.random.default_rng(seed)
    pat = np.where(rng.random((g.ny, g.nx)) < .5, -1., 1.)
    moment = 2e-16 * (1 + 0.2 * rng.random((g.ny, g.nx)))
    return pat, moment


def test_closed_forms_line_and_side():
    a, d = 500e-9, 200e-9
    g = lattices.make("IP_Ising", a, n=7)
    mm = model.OracleMagnets(g, T=300, moment=1e-16, pattern=1, energies=[])
    K, Kps, Kpo, idx = model.build_monopole_kernels(mm, np.full((1, 1), d), 150e-9)
    c = (g.ny - 1, g.nx - 1)
    assert K[0][c] == 0 and Kps[0][c] == 0 and Kpo[0][c] == 0
    for k in range(1, 4):
        r = k * a
        line = 2 / r - 1 / (r + d) - 1 / (r - d)  # head to tail: NN, SS at r; NS, SN at r +- d
        side = 2 / r - 2 / math.hypot(r, d)       # side by side
        # each 1/r term carries the rounding of an FP32 square: at most 2^-24 relative
        tol = 1e-7 * 2 ** -23 * (2 / r + 1 / (r + d) + 1 / (r - d))
        for s in (1, -1):
            assert K[0][c[0], c[1] + s * k] == pytest.approx(1e-7 * line, rel=0, abs=tol)
            assert K[0][c[0] + s * k, c[1]] == pytest.approx(1e-7 * side, rel=0, abs=tol)
            # a magnet turned by 90 degrees on the axis of another: no net interaction
            for Kp in (Kps, Kpo):
                assert abs(Kp[0][c[0], c[1] + s * k]) < 1e-12 * 1e-7 / r
                assert abs(Kp[0][c[0] + s * k, c[1]]) < 1e-12 * 1e-7 / r
    assert K[0][c[0], c[1] + 1] < 0 < K[0][c[0] + 1, c[1]]  # head to tail attracts, side by side repels


@pytest.mark.parametrize("lat,n,pbc", [("IP_Square", 6, False), ("IP_Kagome", 12, False), ("IP_Pinwheel", 6, False),
                                       ("IP_Square", 6, True), ("IP_Pinwheel", 6, True)])
def test_brute_force_energy(lat, n, pbc):
    g = lattices.make(lat, 1e-6, n=n, PBC=pbc)
    pat, moment = _random_state(g, 11)
    dsite_uc = 220e-9 * (1 + 0.1 * np.arange(g.uy * g.ux).reshape(g.uy, g.ux) / (g.uy * g.ux))
    dsite = np.tile(dsite_uc, (math.ceil(g.ny / g.uy), math.ceil(g.nx / g.ux)))[:g.ny, :g.nx]
    dm = model.DiMonopolar(prefactor=1.3, d=dsite_uc, small_d=130e-9)
    mm = model.OracleMagnets(g, T=300, moment=moment, pattern=pat, energies=[dm], PBC=pbc, m_perp_factor=1)
    mm.update_energy()
    Eb, Epb, Eabs = _brute_energy(g, mm.m, moment, dsite, 130e-9, pbc=pbc, pf=1.3)
    # the plain kernel rounds squared distances to FP32 as the reference does: each 1/r term is
    # off by at most 2^-24 of itself
    assert np.all(np.abs(dm.E - Eb) <= 2 ** -23 * Eabs + 1e-12 * np.abs(Eb).max())
    np.testing.assert_allclose(dm.E_perp, Epb, rtol=0, atol=1e-11 * np.abs(Epb).max())


def test_far_field_is_dipolar():
    g = lattices.make("IP_Square", 1e-6, n=8)
    d, mu = 20e-9, 2e-16
    mm = model.OracleMagnets(g, T=300, moment=mu, pattern=1, energies=[])
    K, Kps, Kpo, idx = model.build_monopole_kernels(mm, np.full((g.uy, g.ux), d), d)
    Kd, Kpsd, Kpod, idxd = model.build_kernels(mm)
    assert np.array_equal(idx, idxd)
    q2 = (mu / d) ** 2
    for mono, dip in ((K, Kd), (Kps, Kpsd), (Kpo, Kpod)):
        big = np.abs(dip) > 1e-3 * np.abs(dip).max()
        assert big.sum() > 20
        np.testing.assert_allclose(mono[big] * q2, dip[big] * mu * mu, rtol=1e-2)


@pytest.mark.parametrize("lat,pbc", [("IP_Kagome", False), ("IP_Square", True)])
def test_incremental_equals_full_recompute(lat, pbc):
    """With PBC only the per-switch sums (energies.py:731-736) are exact: the convolution branch
    (energies.py:720-729) wraps a kernel that already holds the periodic images, so the
    reference counts those twice unless REDUCED_KERNEL_SIZE keeps the window below the grid."""
    g = lattices.make(lat, 1e-6, n=12, PBC=pbc)
    pat, moment = _random_state(g, 4)
    dm = model.DiMonopolar(d=200e-9, small_d=150e-9)
    mm = model.OracleMagnets(g, T=300, moment=moment, pattern=pat, energies=[dm, model.Zeeman(1e-3, .3)], PBC=pbc,
                             m_perp_factor=1,
                             params=dict(REDUCED_KERNEL_SIZE=0,
                                         SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=100 if pbc else 3))
    rng = np.random.default_rng(8)
    occ = np.argwhere(g.occupation)
    for it in range(20):
        k = 1 if it % 3 == 0 else int(rng.integers(2, 10))
        sel = occ[rng.choice(len(occ), k, replace=False)].T
        mm.m[sel[0], sel[1]] *= -1
        mm.update_energy(sel)
    E_inc, Ep_inc = dm.E.copy(), dm.E_perp.copy()
    mm.update_energy()
    scale = np.abs(dm.E).max()
    np.testing.assert_allclose(E_inc, dm.E, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(Ep_inc, dm.E_perp, rtol=0, atol=1e-12 * scale)
