"""The oracle replays the reference's own trajectories (golden fixtures) exactly.

The fixtures were written by tests/golden/make_golden.py from the reference implementation
with the Philox streams replayed into it. Switch sequences must match bit-for-bit, energies to
1e-12 relative (both sides are FP64; only summation order can differ).
"""
import numpy as np
import pytest

from cases import CASES, PROGRESS, load, oracle_magnets


@pytest.mark.parametrize("name", sorted(CASES))
def test_oracle_replays_reference(name):
    fx = load(name)
    spec = CASES[name]
    mm = oracle_magnets(name, fx)
    scale = np.max(np.abs(fx["E0"]))
    np.testing.assert_array_equal(mm.m, fx["m0"])
    np.testing.assert_allclose(mm.E, fx["E0"], rtol=0, atol=1e-12 * scale)
    if mm.USE_PERP_ENERGY:
        np.testing.assert_allclose(mm.E_perp, fx["Eperp0"], rtol=0, atol=1e-12 * scale)
    sw_y, sw_x, sw_step = [], [], []
    snaps = list(fx["snap_steps"])
    k = 0
    for s in range(spec["nsteps"]):
        r = mm.update(**spec["update_kw"])
        if r is not None:
            arr = np.asarray(r).reshape(2, -1)
            sw_y += arr[0].tolist()
            sw_x += arr[1].tolist()
            sw_step += [s] * arr.shape[1]
        assert mm.attempted_switches == fx["attempts"][s], f"step {s}"
        assert mm.t == pytest.approx(fx["t"][s], rel=1e-12, abs=0), f"step {s}"
        if k < len(snaps) and s == snaps[k]:
            np.testing.assert_array_equal(mm.m, fx["snap_m"][k])
            np.testing.assert_allclose(mm.E, fx["snap_E"][k], rtol=0, atol=1e-11 * scale)
            if mm.USE_PERP_ENERGY:
                np.testing.assert_allclose(mm.E_perp, fx["snap_Eperp"][k], rtol=0, atol=1e-11 * scale)
            k += 1
    np.testing.assert_array_equal(np.array(sw_y), fx["sw_y"])
    np.testing.assert_array_equal(np.array(sw_x), fx["sw_x"])
    np.testing.assert_array_equal(np.array(sw_step), fx["sw_step"])
    np.testing.assert_array_equal(mm.m, fx["m_final"])
    assert mm.E_tot == pytest.approx(float(fx["E_tot_final"]), rel=1e-10)


def test_golden_specs_match():
    """tests/cases.py mirrors the generator's specs."""
    import os
    src = open(os.path.join(os.path.dirname(__file__), "golden", "make_golden.py")).read()
    for name, spec in list(CASES.items()) + list(PROGRESS.items()):
        fx = load(name)
        text = str(fx["spec"])
        for key in ("lattice", "n", "T", "T_ramp", "moment", "seed", "nsteps", "calls"):
            if key in spec:
                assert repr(spec[key]) in text or str(spec[key]) in text, (name, key)
        assert f'"{name}"' in src


def _count_singles(om):
    """Counts the oracle's _select_single draws (the 'single' scheme, or the grid's fallback)."""
    calls = [0]
    orig = om._select_single

    def single(attempt):
        calls[0] += 1
        return orig(attempt)

    om._select_single = single
    return calls


@pytest.mark.parametrize("name", ["metro_kagome_single", "metro_oop_fallback"])
def test_oracle_single_draws_like_the_reference(name):
    """_select_single (core.py:489-494) is drawn as often as in the reference: every step of the
    'single' scheme, and each time the grid's samples all fall outside the grid (core.py:555-558)
    -- which the fallback fixture does in more than half of its steps."""
    fx = load(name)
    spec = CASES[name]
    om = oracle_magnets(name, fx)
    calls = _count_singles(om)
    for _ in range(spec["nsteps"]):
        om.update(**spec["update_kw"])
    assert calls[0] == int(fx["singles"])
    assert calls[0] > (spec["nsteps"] // 2 if name.endswith("fallback") else spec["nsteps"] - 1)
    np.testing.assert_array_equal(om.m, fx["m_final"])


def _progress_log(om):
    """Logs the switches of every update() that progress() makes (through self.update)."""
    log = dict(y=[], x=[], step=[])
    orig = om.update

    def update(*args, **kwargs):
        step = om.step
        r = orig(*args, **kwargs)
        if r is not None:
            arr = np.asarray(r).reshape(2, -1)
            log["y"] += arr[0].tolist()
            log["x"] += arr[1].tolist()
            log["step"] += [step] * arr.shape[1]
        return r

    om.update = update
    return log


@pytest.mark.parametrize("name", sorted(PROGRESS))
def test_oracle_replays_reference_progress(name):
    """Magnets.progress (core.py:736-760) call by call: the returned (dt, dMCsteps), the number of
    update() steps each call makes (both stop predicates, rtol 1e-10, and the t_max - dt that
    Neel steps get), every switch, and the spins after each call."""
    fx = load(name)
    spec = PROGRESS[name]
    om = oracle_magnets(name, fx)
    log = _progress_log(om)
    scale = np.max(np.abs(fx["E0"]))
    k = 0
    snaps = list(fx["snap_steps"])
    for c, call in enumerate(spec["calls"]):
        dt, dmc = om.progress(**call)
        assert om.step == fx["steps"][c], f"call {c}: {om.step} update() steps, reference {fx['steps'][c]}"
        assert om.attempted_switches == fx["attempts"][c] and om.switches == fx["switches_hist"][c]
        assert dt == pytest.approx(fx["ret"][c][0], rel=1e-12, abs=0), f"call {c}"
        assert dmc == pytest.approx(fx["ret"][c][1], rel=1e-12, abs=0), f"call {c}"
        assert om.t == pytest.approx(fx["t"][c], rel=1e-12, abs=0)
        np.testing.assert_array_equal(om.m, fx["m_hist"][c])
        if k < len(snaps) and c == snaps[k]:
            np.testing.assert_allclose(om.E, fx["snap_E"][k], rtol=0, atol=1e-11 * scale)
            k += 1
    np.testing.assert_array_equal(np.array(log["y"]), fx["sw_y"])
    np.testing.assert_array_equal(np.array(log["x"]), fx["sw_x"])
    np.testing.assert_array_equal(np.array(log["step"]), fx["sw_step"])
    assert om.E_tot == pytest.approx(float(fx["E_tot_final"]), rel=1e-10)


def test_progress_fixtures_reach_each_stop():
    """The progress fixtures exercise what they are for: Neel calls ended by MCsteps and by time,
    a time-bound call whose last step got the remaining t_max - dt and did not switch
    (core.py:701-704, 757), a one-step call with no switch; Metropolis calls ended by MCsteps
    and, with METROPOLIS_TIME, by time."""
    from oracle.model import lower_than
    fx = load("prog_neel_kagome")
    calls = PROGRESS["prog_neel_kagome"]["calls"]
    steps = np.diff(np.concatenate([[0], fx["steps"]]))
    sws = np.diff(np.concatenate([[0], fx["switches_hist"]]))
    by_time = [c for c, call in enumerate(calls) if not lower_than(fx["ret"][c][0], call["t_max"])]
    by_mc = [c for c, call in enumerate(calls) if not lower_than(fx["ret"][c][1], call["MCsteps_max"])]
    assert by_time and by_mc
    assert any(sws[c] == steps[c] - 1 and steps[c] > 1 for c in by_time)  # the hand-off step
    assert any(steps[c] == 1 and sws[c] == 0 for c in by_time)
    fx = load("prog_metro_time")
    calls = PROGRESS["prog_metro_time"]["calls"]
    assert any(not lower_than(fx["ret"][c][0], call["t_max"]) for c, call in enumerate(calls))
    assert any(not lower_than(fx["ret"][c][1], call["MCsteps_max"]) for c, call in enumerate(calls))


def test_progress_fixture_catches_a_missing_hand_off():
    """A planted mistake in the oracle: Neel steps inside progress() get t_max instead of the
    remaining t_max - dt (core.py:757). The fixture's time-bound calls then diverge from the
    reference, so the fixture pins that term."""
    import math
    from oracle.model import lower_than
    fx = load("prog_neel_kagome")
    om = oracle_magnets("prog_neel_kagome", fx)

    def progress_without_hand_off(t_max=1, MCsteps_max=4.0):
        t0, mc0 = om.t, om.MCsteps
        while lower_than(om.t - t0, t_max) and lower_than(om.MCsteps - mc0, MCsteps_max):
            om.update(t_max=t_max)
        return om.t - t0, om.MCsteps - mc0

    diverged = False
    for c, call in enumerate(PROGRESS["prog_neel_kagome"]["calls"]):
        dt, _ = progress_without_hand_off(**call)
        if om.switches != fx["switches_hist"][c] or not math.isclose(dt, fx["ret"][c][0], rel_tol=1e-12):
            diverged = True
            break
    assert diverged
