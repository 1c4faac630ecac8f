"""Metropolis steps over the caller's own samples, update(idx=...) and progress(idx=...)
(core.py:659-661, 736-760), against the oracle: same accepted cells, spins, counters, fields
and (METROPOLIS_TIME) elapsed time. Acceptance uniforms are keyed by cell and step, so the
oracle draws the same ones; a sample whose uniform lies within NEAR_TIE (relative) of its
probability may differ (FP32 fields vs FP64), and the oracle then continues from the device."""
import math

import numpy as np
import pytest

from cases import energy_scale, load, oracle_magnets, product_magnets

pytestmark = pytest.mark.gpu

NEAR_TIE = 1e-4
E_TOL = 1e-5


def _pick(rng, mm, k):
    occ = np.flatnonzero(np.asarray(mm.occupation).reshape(-1))
    flat = rng.choice(occ, size=min(k, occ.size), replace=False)
    return np.asarray([flat // mm.nx, flat % mm.nx])


def _same_or_tie(om, a, b):
    if a.shape == b.shape and np.array_equal(a, b):
        return True
    idx, _, ua, expo = om.last_samples
    close = np.abs(ua - expo) <= NEAR_TIE * expo
    near = set((idx[0][close] * om.nx + idx[1][close]).tolist())
    diff = set((a[0] * om.nx + a[1]).tolist()) ^ set((b[0] * om.nx + b[1]).tolist())
    assert diff and diff <= near, "accepted sets differ outside the near-tie samples"
    return False


@pytest.mark.parametrize("name", ["metro_oop_conv", "metro_pinwheel", "metro_ipsq_pbc", "metro_oop_window_large"])
def test_update_with_explicit_samples_matches_oracle(name):
    fx = load(name)
    mm = product_magnets(name, fx)
    om = oracle_magnets(name, fx)
    scale = energy_scale(om)
    rng = np.random.default_rng(7)
    for s, k in enumerate([1, 5, 40, 0, 200, 3, 64, 1000, 17]):
        idx = _pick(rng, mm, k) if k else np.zeros((2, 0), dtype=np.int64)
        a = np.asarray(mm.update(idx=idx)).reshape(2, -1)
        b = np.asarray(om.update(idx=idx)).reshape(2, -1)
        if not _same_or_tie(om, a, b):
            om.m[...] = np.asarray(mm.m, dtype=float)
            om.update_energy()
            om.t, om.switches = float(mm.t), int(mm.switches)
            continue
        assert mm.attempted_switches == om.attempted_switches
        assert mm.switches == om.switches
        np.testing.assert_array_equal(mm.m, om.m)
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=E_TOL * scale, err_msg=f"{name} step {s}")
        if om.t != 0:
            assert mm.t == pytest.approx(om.t, rel=1e-4)


def test_progress_with_explicit_samples_matches_oracle():
    name = "metro_oop_conv"
    fx = load(name)
    mm = product_magnets(name, fx)
    om = oracle_magnets(name, fx)
    idx = _pick(np.random.default_rng(3), mm, 96)
    a = mm.progress(t_max=math.inf, MCsteps_max=1.5, idx=idx)
    b = om.progress(t_max=math.inf, MCsteps_max=1.5, idx=idx)
    assert mm.attempted_switches == om.attempted_switches
    assert mm.switches == om.switches
    np.testing.assert_array_equal(mm.m, om.m)
    assert a[1] == pytest.approx(b[1], rel=1e-12)


def test_batched_explicit_samples_equal_alone():
    from hotspice_amd import batch
    names = ["metro_oop_conv", "metro_oop_window_large"]

    def make():
        return [product_magnets(n, load(n)) for n in names]

    alone = make()
    kws = [dict(idx=_pick(np.random.default_rng(i), m, 50)) for i, m in enumerate(alone)]
    for m, kw in zip(alone, kws):
        m.progress(t_max=math.inf, MCsteps_max=0.5, **kw)
    many = make()
    batch.progress_many(many, t_max=math.inf, MCsteps_max=0.5, per_instance=kws)
    for m1, m2 in zip(alone, many):
        np.testing.assert_array_equal(m1.m, m2.m)
        np.testing.assert_array_equal(np.asarray(m1.E), np.asarray(m2.E))
        assert (m1.attempted_switches, m1.switches, m1.t) == (m2.attempted_switches, m2.switches, m2.t)


def test_explicit_samples_are_checked():
    mm = product_magnets("metro_oop_conv", load("metro_oop_conv"))
    with pytest.raises(ValueError):
        mm.update(idx=np.array([[0, 0], [1, 1]]))  # a position twice
    with pytest.raises(ValueError):
        mm.update(idx=np.array([[mm.ny], [0]]))  # outside the grid
    with pytest.raises(ValueError):
        mm.update(idx=np.array([0.5, 1.0]))  # not (2, k) integers


def test_device_rejects_a_bad_sample_list():
    """Through the C ABI a list may hold anything: a cell outside the grid or without a magnet ends
    the step with err = 4 before anything changes (the kernel checks every listed cell)."""
    import torch
    from hotspice_amd import _lib
    mm = product_magnets("metro_oop_window_large", load("metro_oop_window_large"))
    mm._pre_step()
    m0 = mm._dev.m.clone()
    E0 = mm._dev.E.clone()
    for bad in ([5, -1, 7], [5, mm.nx * mm.ny, 7]):
        sel = _lib.Select()
        sel.scheme, sel.nsc = 3, len(bad)
        cells = torch.tensor(bad, dtype=torch.int32, device=mm._dev.device)
        sel.pat_cells = cells.data_ptr()
        run = _lib.Run(t_max=math.inf, mc_max=math.inf, attempt_freq=1e10, max_steps=1, log=0, nblocks=0, begin=1,
                       first_free=1)
        mm._dev.metropolis_run(sel, run)
        st = mm._state()
        assert st.err == 4 and st.attempts == 0
        assert torch.equal(mm._dev.m, m0) and torch.equal(mm._dev.E, E0)
