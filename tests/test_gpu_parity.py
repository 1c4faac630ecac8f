"""Parity of the HIP stepping path with the FP64 oracle and the reference's golden trajectories.

Bars (written in the tests):
  * integer/index results (Philox words -> uniforms, switch indices, accepted sets, spins) are
    bit-exact, up to the first FP32-vs-FP64 near-tie, which is reported: a Neel step whose two
    smallest waiting times differ by < NEAR_TIE relative, or a Metropolis sample whose
    acceptance uniform lies within NEAR_TIE (relative) of its probability;
  * energies within 1e-5 relative to the field scale (FP32 + Kahan fields vs FP64; the scale is
    the largest |E| or, where terms cancel, the size of one exchange or Zeeman term), initial
    fields (FP64 recompute stored as FP32 + compensation) within 1e-9;
  * elapsed time within 1e-5 relative, plus what the energy error implies: a Neel step lasts
    exp(beta*barrier_w)*(-ln u)/f (the global minimum barrier cancels), and a barrier moves by at
    most 2*max|dE| (+ max|dE_perp|), so each step's time may differ by expm1(2*beta*that)
    relative; this grows with beta*|E|, which reaches ~100 for closely spaced dumbbell charges.
    E_B is held in FP32 (2^-24 relative), which adds beta*E_B*2^-24 relative per step: 1.7e-5 at
    a 290 kT barrier. Metropolis steps with METROPOLIS_TIME get the same bound.
    A step time that underflows into the subnormals (~1e-320 s in a deep minimum) has no
    relative precision left: the device clears the low 6 bits of every step time (so that all
    stepping paths add the same value), which there is up to 63 subnormal units, so each step
    adds 64 units of absolute slack.
"""
import math

import numpy as np
import pytest

from cases import CASES, PROGRESS, SYNTH, energy_scale, load, oracle_magnets, product_magnets, spec as case_spec

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-4
E_TOL = 1e-5


def _neel_near_tie(om, g_idx, o_idx):
    taus = om.last_taus
    tg, to = taus[divmod(g_idx, om.nx)], taus[divmod(o_idx, om.nx)]
    return abs(tg / to - 1) < NEAR_TIE


def _metropolis_near_tie(om, a, b):
    """True when the device's accepted cells `a` and the oracle's `b` differ only in samples whose
    acceptance uniform lies within NEAR_TIE (relative) of their probability: every cell of the
    symmetric difference must be such a sample, and the difference must not be empty (equal sets
    listed in another order are a real mismatch)."""
    idx, _, ua, expo = om.last_samples
    close = np.abs(ua - expo) <= NEAR_TIE * expo
    near = set((idx[0][close] * om.nx + idx[1][close]).tolist())
    ca = set((a[0] * om.nx + a[1]).tolist()) if a.size else set()
    cb = set((b[0] * om.nx + b[1]).tolist()) if b.size else set()
    diff = ca ^ cb
    return bool(diff) and diff <= near


def _continue_from_device(om, mm):
    """After a near-tie the oracle continues from the device's state: its spins, time and
    counters, with its FP64 fields recomputed from those spins."""
    om.m[...] = np.asarray(mm.m, dtype=float)
    om.update_energy()
    om.t = float(mm.t)
    om.switches = int(mm.switches)
    om.attempted_switches = int(mm.attempted_switches)


def lockstep(name, nsteps=None, check_every=25, min_compared=0.5, check=None):
    """Steps product and oracle side by side; returns (steps compared, near-tie steps). At a
    near-tie the oracle continues from the device's state, so the steps after it are compared
    too; fewer than `min_compared` of the steps compared without a near-tie fails. `check(mm,
    recomputes)` runs at the end with the FP64 recomputes the stepping ran."""
    spec = case_spec(name)
    fx = load(name)
    mm = product_magnets(name, fx, **spec.get("product_kw", {}))
    om = oracle_magnets(name, fx)
    recomputes0 = mm.recomputes
    scale = energy_scale(om)
    neel = spec["params"]["UPDATE_SCHEME"] in ("Néel", "Neel")
    timed = neel or bool(spec["params"].get("METROPOLIS_TIME", False))  # times exp(barrier * beta)
    nsteps = spec["nsteps"] if nsteps is None else nsteps
    t_slack = 0.0
    beta_EB = float(np.max(np.abs(om.E_B) * om.beta))  # E_B is held in FP32 on the device
    ties = []
    for s in range(nsteps):
        if timed:
            dE = float(np.max(np.abs(np.asarray(mm.E, dtype=float) - om.E)))
            if om.USE_PERP_ENERGY:
                dE += float(np.max(np.abs(np.asarray(mm.E_perp, dtype=float) - om.E_perp)))
            t_0 = om.t
        a = mm.update(**spec["update_kw"])
        b = om.update(**spec["update_kw"])
        if timed and math.isfinite(om.t - t_0):
            t_slack += abs(om.t - t_0) * math.expm1(2 * float(np.max(om.beta)) * dE + beta_EB * 2.0 ** -24)
        tie = False
        if neel:
            ga = -1 if a is None else int(a[0]) * om.nx + int(a[1])
            gb = -1 if b is None else int(b[0]) * om.nx + int(b[1])
            if ga != gb:
                assert ga >= 0 and gb >= 0 and _neel_near_tie(om, ga, gb), f"{name}: Neel step {s}: {a} vs {b}"
                tie = True
        else:
            a, b = np.asarray(a), np.asarray(b)
            if a.shape != b.shape or not np.array_equal(a, b):
                assert _metropolis_near_tie(om, a.reshape(2, -1), b.reshape(2, -1)), \
                    f"{name}: Metropolis step {s} differs outside the near-tie samples"
                tie = True
        if tie:
            print(f"{name}: FP32-vs-FP64 near-tie at step {s}; continuing from the device state")
            ties.append(s)
            _continue_from_device(om, mm)
            continue
        assert mm.attempted_switches == om.attempted_switches
        assert mm.switches == om.switches
        if om.t != 0 or mm.t != 0:
            t_tol = E_TOL * abs(om.t) + t_slack + (s + 1) * 64 * math.ulp(0.0)
            # equal values pass as such: barriers of ~450 kT overflow both times to +inf alike
            assert mm.t == om.t or abs(mm.t - om.t) <= t_tol, f"{name}: t at step {s}: {mm.t} vs {om.t}"
        if (s + 1) % check_every == 0:
            np.testing.assert_array_equal(mm.m, om.m)
            np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=E_TOL * scale, err_msg=f"{name} step {s}")
            if om.USE_PERP_ENERGY:
                np.testing.assert_allclose(mm.E_perp, om.E_perp, rtol=0, atol=E_TOL * scale)
    np.testing.assert_array_equal(mm.m, om.m)
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=E_TOL * scale, err_msg=f"{name} final")
    assert mm.E_tot == pytest.approx(om.E_tot, rel=E_TOL, abs=E_TOL * scale)
    compared = nsteps - len(ties)
    assert compared >= min_compared * nsteps, f"{name}: only {compared} of {nsteps} steps compared ({ties})"
    if check is not None:
        check(mm, mm.recomputes - recomputes0)
    return compared, ties


@pytest.mark.parametrize("name", sorted(CASES))
def test_initial_fields_match_golden(name):
    fx = load(name)
    mm = product_magnets(name, fx)
    scale = np.abs(fx["E0"]).max()
    np.testing.assert_array_equal(mm.m, fx["m0"])
    np.testing.assert_allclose(mm.E, fx["E0"], rtol=0, atol=1e-9 * scale)
    if mm.USE_PERP_ENERGY:
        np.testing.assert_allclose(mm.E_perp, fx["Eperp0"], rtol=0, atol=1e-9 * scale)
    assert mm.E_tot == pytest.approx(float(fx["E_tot0"]), rel=1e-9, abs=1e-9 * scale)


@pytest.mark.parametrize("name", sorted(CASES) + sorted(SYNTH))
def test_trajectory_matches_oracle(name):
    compared, ties = lockstep(name)
    assert compared >= 1


@pytest.mark.parametrize("name", sorted(CASES))
def test_trajectory_matches_golden(name):
    """The device trajectory equals the reference's own (replayed) trajectory."""
    spec = CASES[name]
    fx = load(name)
    mm = product_magnets(name, fx)
    ys, xs, ss = [], [], []
    for s in range(spec["nsteps"]):
        r = mm.update(**spec["update_kw"])
        if r is not None:
            arr = np.asarray(r).reshape(2, -1)
            ys += arr[0].tolist()
            xs += arr[1].tolist()
            ss += [s] * arr.shape[1]
        assert mm.attempted_switches == fx["attempts"][s]
    k = min(len(ys), len(fx["sw_y"]))
    same = (np.array(ys[:k]) == fx["sw_y"][:k]) & (np.array(xs[:k]) == fx["sw_x"][:k])
    first = int(np.argmin(same)) if not same.all() else k
    assert same.all() and len(ys) == len(fx["sw_y"]), f"{name}: trajectories part at switch {first}"
    np.testing.assert_array_equal(mm.m, fx["m_final"])
    scale = np.abs(fx["E0"]).max()
    np.testing.assert_allclose(mm.E, fx["E_final"], rtol=0, atol=E_TOL * scale)
    assert mm.t == pytest.approx(float(fx["t"][-1]), rel=E_TOL)


@pytest.mark.parametrize("name", sorted(PROGRESS))
def test_progress_matches_golden(name):
    """Device progress() (the step loop with its stop predicates on the device) against the
    reference's own progress() calls, replayed: after every call the same update() steps and
    switches, bit-exact spins, (dt, dMCsteps) as returned. dt within 1e-5 relative plus what the
    FP32 arithmetic implies for step times exp(beta * barrier): beta * max E_B * 2^-24 (FP32 E_B)
    plus 2 * beta * max|E| * E_TOL (the field error bar)."""
    spec = PROGRESS[name]
    fx = load(name)
    mm = product_magnets(name, fx)
    om = oracle_magnets(name, fx)
    beta = float(np.max(om.beta))
    t_rel = E_TOL + beta * float(np.max(np.abs(om.E_B))) * 2.0 ** -24 + 2 * beta * energy_scale(om) * E_TOL
    scale = energy_scale(om)
    for c, call in enumerate(spec["calls"]):
        dt, dmc = mm.progress(**call)
        assert mm.attempted_switches == fx["attempts"][c], f"{name} call {c}"
        assert mm.switches == fx["switches_hist"][c], f"{name} call {c}"
        np.testing.assert_array_equal(mm.m, fx["m_hist"][c], err_msg=f"{name} call {c}")
        assert dmc == pytest.approx(fx["ret"][c][1], rel=1e-12), f"{name} call {c}"
        assert dt == pytest.approx(fx["ret"][c][0], rel=t_rel, abs=0), f"{name} call {c}"
        assert mm.t == pytest.approx(fx["t"][c], rel=t_rel, abs=0), f"{name} call {c}"
    np.testing.assert_allclose(mm.E, fx["E_final"], rtol=0, atol=E_TOL * scale)


def test_philox_device_streams_equal_oracle():
    from oracle import philox
    mm = product_magnets("neel_oop_parabolic")
    d = mm._dev
    for seed, step, tag, i0, n, oz in [(0, 0, 1, 0, 1000, 1), (2 ** 63 + 12345, 7, 3 | (5 << 8), 17, 333, 0),
                                       (99, 2 ** 33 + 1, 4, 10 ** 6, 64, 0)]:
        dev = d.philox_uniform53(seed, step, tag, i0, n, oz)
        ref = philox.uniform53(seed, step, tag, np.arange(i0, i0 + n), open_zero=bool(oz))
        np.testing.assert_array_equal(dev, ref)


@pytest.mark.parametrize("lat,n,kw", [("IP_Kagome", 20, {}), ("IP_Pinwheel", 16, dict(PBC=True)),
                                      ("OOP_Square", 12, dict(PBC=True)), ("IP_Square", 16, {})])
@pytest.mark.parametrize("cutoff", [math.inf, 3.1e-6])
def test_dipolar_kernels_equal_oracle(lat, n, kw, cutoff):
    import hotspice_amd as hs
    from oracle import model
    cls = getattr(hs.ASI, lat)
    mm = cls(1e-6, n, T=300, moment=2e-16, energies=[hs.DipolarEnergy(cutoff_range=cutoff)], seed=1,
             major_axis=200e-9, minor_axis=80e-9, **kw)
    om = model.make_magnets(lat, 1e-6, n=n, T=300, moment=2e-16, pattern=mm.m, energies=[], major_axis=200e-9,
                            minor_axis=80e-9, PBC=kw.get("PBC", False))
    K, Kps, Kpo, idx = model.build_kernels(om, cutoff_range=cutoff)
    d = mm.get_energy('dipolar')
    scale = np.abs(K).max()
    np.testing.assert_allclose(d.kernel_unitcell, K, rtol=0, atol=1e-13 * scale)
    if mm.in_plane:
        np.testing.assert_allclose(d.kernel_perpself_unitcell, Kps, rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(d.kernel_perpother_unitcell, Kpo, rtol=0, atol=1e-13 * scale)
    np.testing.assert_array_equal(d.kernel_unitcell_indices, idx)


@pytest.mark.parametrize("lat,n,kw", [("IP_Kagome", 24, {}), ("IP_Pinwheel", 16, dict(PBC=True)),
                                      ("IP_Square", 14, {}), ("IP_Square", 12, dict(PBC=True))])
def test_dimonopolar_kernels_equal_oracle(lat, n, kw):
    """hs_build_dimonopolar_kernels against the oracle's restatement of energies.py:416-641. The
    plain kernel rounds squared distances to FP32 and takes their FP32 power -0.5 (as the
    reference does); both sides take the correctly rounded FP32 value, so each 1/r term is the
    same FP32 number and the entries (FP64 sums of four terms) agree to FP64 rounding, also
    where the four terms nearly cancel. The closest-charge bound (2^-22 of the largest 1/r term)
    is what a single misrounded FP32 term would cost; the test asserts far below it."""
    import hotspice_amd as hs
    from oracle import model
    cls = getattr(hs.ASI, lat)
    probe = cls(1e-6, n, T=300, moment=2e-16, energies=[], seed=1, **kw)
    d_uc = 300e-9 * (1 + 0.05 * np.arange(probe.unitcell.y * probe.unitcell.x).reshape(probe.unitcell.y, probe.unitcell.x))
    dm = hs.energies.DiMonopolarEnergy(d=d_uc, small_d=120e-9)
    mm = cls(1e-6, n, T=300, moment=2e-16, energies=[dm], seed=1, **kw)
    om = model.make_magnets(lat, 1e-6, n=n, T=300, moment=2e-16, pattern=mm.m, energies=[], PBC=kw.get("PBC", False))
    K, Kps, Kpo, idx = model.build_monopole_kernels(om, d_uc, 120e-9)
    np.testing.assert_array_equal(dm.kernel_unitcell_indices, idx)
    ys, xs = np.nonzero(om.occupation)
    ang = om.geom.angles[ys, xs]
    h = np.tile(d_uc, (math.ceil(om.ny / om.uy), math.ceil(om.nx / om.ux)))[ys, xs] / 2
    qx = np.concatenate([om.x[xs] + h * np.cos(ang), om.x[xs] - h * np.cos(ang)])
    qy = np.concatenate([om.y[ys] + h * np.sin(ang), om.y[ys] - h * np.sin(ang)])
    owner = np.concatenate([np.arange(len(ys))] * 2)
    dist = np.hypot(qx[:, None] - qx[None, :], qy[:, None] - qy[None, :])
    rmin = dist[owner[:, None] != owner[None, :]].min()  # closest charges of two magnets
    one_term_ulp = 1e-7 * 2.0 ** -22 / rmin
    assert np.abs(dm.kernel_unitcell - K).max() < 1e-3 * one_term_ulp
    np.testing.assert_allclose(dm.kernel_unitcell, K, rtol=0, atol=1e-13 * np.abs(K).max())
    scale = np.abs(Kps).max()
    np.testing.assert_allclose(dm.kernel_perpself_unitcell, Kps, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(dm.kernel_perpother_unitcell, Kpo, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("lat,n,kw", [("IP_Kagome", 24, {}), ("IP_Square", 12, dict(PBC=True))])
def test_dimonopolar_fields_equal_oracle(lat, n, kw):
    """FP64 recompute with DiMonopolarEnergy: non-uniform moments (charge q = moment/d of the
    receiving magnet, energies.py:654-659), Zeeman from the moments, E and E_perp."""
    import hotspice_amd as hs
    from oracle import model
    cls = getattr(hs.ASI, lat)
    probe = cls(1e-6, n, T=300, moment=2e-16, energies=[], seed=1, **kw)
    rng = np.random.default_rng(3)
    moment = 2e-16 * (1 + 0.3 * rng.random(probe.m.shape))
    pat = np.where(rng.random(probe.m.shape) < .5, -1., 1.)
    mm = cls(1e-6, n, T=300, moment=moment, pattern=pat, seed=1, m_perp_factor=0.7,
             energies=[hs.energies.DiMonopolarEnergy(prefactor=0.8, d=280e-9, small_d=140e-9),
                       hs.ZeemanEnergy(2e-3, 0.6)], **kw)
    om = model.make_magnets(lat, 1e-6, n=n, T=300, moment=moment, pattern=mm.m, m_perp_factor=0.7,
                            energies=[model.DiMonopolar(prefactor=0.8, d=280e-9, small_d=140e-9), model.Zeeman(2e-3, 0.6)],
                            PBC=kw.get("PBC", False))
    om.update_energy()
    scale = np.abs(om.E).max()
    np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-9 * scale)
    np.testing.assert_allclose(mm.E_perp, om.E_perp, rtol=0, atol=1e-9 * scale)
    assert mm.E_tot == pytest.approx(om.E_tot, rel=1e-9, abs=1e-9 * scale)
    d = mm.get_energy('dimonopolar')
    np.testing.assert_allclose(d.E, om.get_energy('dimonopolar').E, rtol=0, atol=1e-9 * scale)


@pytest.mark.parametrize("name", ["neel_pinwheel_tramp", "neel_oop_tramp_parabolic"])
def test_progress_varying_temperature_matches_oracle(name):
    """progress() with a temperature that varies over the grid (two launches per step on the
    device, core.py:686-699 looped as core.py:736-760) against the oracle's progress()."""
    fx = load(name)
    mm = product_magnets(name, fx)
    om = oracle_magnets(name, fx)
    assert not mm._uniform_beta
    for kw in (dict(t_max=math.inf, MCsteps_max=0.3), dict(t_max=1e-3, MCsteps_max=0.2)):
        a = mm.progress(**kw)
        b = om.progress(**kw)
        assert mm.attempted_switches == om.attempted_switches
        assert mm.switches == om.switches
        np.testing.assert_array_equal(mm.m, om.m)
        assert a[1] == pytest.approx(b[1], rel=1e-12)
        assert mm.t == pytest.approx(om.t, rel=1e-4)
