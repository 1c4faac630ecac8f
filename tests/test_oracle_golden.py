"""The oracle replays the reference's own trajectories (golden fixtures) exactly.

The fixtures were written by tests/golden/make_golden.py from the reference implementation
with the Philox streams replayed into it. Switch sequences must match bit-for-bit, energies to
1e-12 relative (both sides are FP64; only summation order can differ).
"""
import numpy as np
import pytest

from cases import CASES, load, oracle_magnets


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
    import ast
    import os
    src = open(os.path.join(os.path.dirname(__file__), "golden", "make_golden.py")).read()
    for name, spec in CASES.items():
        fx = load(name)
        text = str(fx["spec"])
        for key in ("lattice", "n", "T", "moment", "seed", "nsteps"):
            assert repr(spec[key]) in text or str(spec[key]) in text, (name, key)
        assert f'"{name}"' in src
