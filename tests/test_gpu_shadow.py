"""Periodic FP64 recomputes refreshed from the FP64 shadow of the dipolar correlations
(hs_refresh_fields): the raw correlations sum_j K[j - i] (m w)_j of the last full recompute plus
the change (m - m_ref) w of every cell that switched since, over its window, must give the
fields of a fresh full recompute of the same spins to FP64 rounding (1e-12 of the field scale),
for E and E_perp, for moments and dumbbell charges, truncated and full windows, FP32 + Kahan
and FP64 fields. The changed cells are added in cell order, so a refresh is deterministic; a
change of weights, occupation or kernels makes the next periodic recompute a full one."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

kB = 1.380649e-23


def _inputs(shape, seed, T=300, EB_kT=20):
    rng = np.random.default_rng(seed)
    return dict(E_B=EB_kT * kB * T * rng.normal(1, 0.05, size=shape), pattern=np.where(rng.random(shape) < .5, -1., 1.))


def _build(kind, seed=1):
    """(instance, step function advancing a few switches)."""
    import hotspice_amd as hs
    if kind == "neel_pinwheel_perp":  # full windows, E and E_perp, Zeeman at an angle
        n = 64
        mm = hs.ASI.IP_Pinwheel(1e-6, n, T=300, moment=2.37e-16, seed=seed, recompute_every=None,
                                energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3, math.pi / 8)],
                                params=hs.SimParams(UPDATE_SCHEME='Néel'), **_inputs((n, n), seed, EB_kT=40))
        return mm, lambda: mm.update()
    if kind == "neel_kagome_mono":  # dumbbell kernels, unit sources times charges
        mm = hs.ASI.IP_Kagome(1e-6, 40, T=300, moment=2.37e-16, seed=seed, recompute_every=None,
                              energies=[hs.energies.DiMonopolarEnergy(d=350e-9, small_d=150e-9),
                                        hs.ZeemanEnergy(5e-4, 0.3)], params=hs.SimParams(UPDATE_SCHEME='Néel'))
        mm.E_B = 40 * kB * 300 * np.random.default_rng(seed).normal(1, 0.05, size=mm.shape)
        return mm, lambda: mm.update()
    if kind == "metro_ipsq_perp":  # Metropolis, in-plane with E_perp, full windows
        probe = hs.ASI.IP_Square(1e-6, 32, moment=2.37e-16, energies=[], params=hs.SimParams(UPDATE_SCHEME='Metropolis'))
        shape = (probe.ny, probe.nx)
        del probe
        mm = hs.ASI.IP_Square(1e-6, 32, T=300, moment=2.37e-16, seed=seed, recompute_every=None,
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(2e-3, 0.7)],
                              params=hs.SimParams(UPDATE_SCHEME='Metropolis'), **_inputs(shape, seed))
        return mm, lambda: mm.update(r=8e-6)
    if kind == "metro_oop_window_exchange":  # truncated windows, exchange and a field
        n = 96
        kT = kB * 500
        mm = hs.ASI.OOP_Square(1e-6, n, T=500, moment=1.1e-15, seed=seed, recompute_every=None,
                               energies=[hs.DipolarEnergy(cutoff_range=5e-6), hs.ZeemanEnergy(1e-3),
                                         hs.ExchangeEnergy(J=0.3 * kT)],
                               params=hs.SimParams(UPDATE_SCHEME='Metropolis', REDUCED_KERNEL_SIZE=5),
                               **_inputs((n, n), seed, T=500))
        return mm, lambda: mm.update(r=12e-6)
    # FP64 fields (the 64x64 plain configuration)
    from hotspice_amd.workloads import oop_metropolis
    mm = oop_metropolis(fp64=True, seed=seed)
    return mm, lambda: mm.update(r=8 * 200e-9)


def _fields(mm):
    return mm.E.copy(), (mm.E_perp.copy() if mm.USE_PERP_ENERGY else None)


def _step_some(mm, step, lo=10):
    """Steps until at least `lo` switches and at most a sixteenth of the grid have happened since
    the last recompute (a refresh is then due at the next periodic recompute)."""
    cap = mm.nx * mm.ny // 16
    for _ in range(2000):
        bound = mm._changed_bound()
        if bound is not None and bound >= lo:
            break
        step()
    bound = mm._changed_bound()
    assert bound is not None and lo <= bound <= cap, (bound, cap)


KINDS = ["neel_pinwheel_perp", "neel_kagome_mono", "metro_ipsq_perp", "metro_oop_window_exchange", "fp64_oop"]


@pytest.mark.parametrize("kind", KINDS)
def test_refresh_equals_full_recompute(kind):
    mm, step = _build(kind)
    if mm.params.UPDATE_SCHEME != 'Néel':  # the Metropolis bound counts switches from a recompute on
        mm._device_recompute(periodic=True)
    for rnd in range(3):
        _step_some(mm, step)
        r0 = mm._dev.shadow_refreshes
        m_before = mm.m.copy()
        mm._device_recompute(periodic=True)
        assert mm._dev.shadow_refreshes == r0 + 1, "the refresh path was not taken"
        assert np.array_equal(mm.m, m_before)
        E_r, P_r = _fields(mm)
        saved = mm._dev.shadow_refreshes
        mm._device_recompute()  # full
        assert mm._dev.shadow_refreshes == 0
        E_f, P_f = _fields(mm)
        scale = float(np.max(np.abs(E_f)))
        err = float(np.max(np.abs(E_r - E_f))) / scale
        assert err < 1e-12, f"{kind} round {rnd}: E differs from the full recompute by {err:.3g} of its scale"
        if P_f is not None:
            errp = float(np.max(np.abs(P_r - P_f))) / max(float(np.max(np.abs(P_f))), 1e-300)
            assert errp < 1e-12, f"{kind} round {rnd}: E_perp differs by {errp:.3g}"
        assert saved >= 1


def test_refresh_is_deterministic():
    """Two instances with one seed stepped alike: their refreshed fields agree bit for bit."""
    out = []
    for _ in range(2):
        mm, step = _build("neel_pinwheel_perp", seed=7)
        _step_some(mm, step, lo=150)
        mm._device_recompute(periodic=True)
        assert mm._dev.shadow_refreshes == 1
        c = mm._dev.components()
        out.append((mm.m.copy(), np.asarray(c[4]).copy(), np.asarray(c[5]).copy()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1]) and np.array_equal(out[0][2], out[1][2])


def test_weight_and_occupation_changes_force_a_full_recompute():
    mm, step = _build("neel_pinwheel_perp", seed=3)
    _step_some(mm, step)
    mm.moment = mm.moment * 1.25  # the shadow holds correlations of the old moments
    mm._device_recompute(periodic=True)
    assert mm._dev.shadow_refreshes == 0
    E_p, _ = _fields(mm)
    mm._device_recompute()
    assert np.array_equal(E_p, mm.E)
    _step_some(mm, step)
    mm._device_recompute(periodic=True)
    assert mm._dev.shadow_refreshes == 1  # the full recompute stored a new shadow
    # spins written from the host between recomputes are changes like any other
    m = mm.m.copy()
    m[5:9, 5:9] *= -1
    mm.m = m
    mm._sync_m()
    mm._device_recompute(periodic=True)
    assert mm._dev.shadow_refreshes == 2
    E_r, P_r = _fields(mm)
    mm._device_recompute()
    E_f, P_f = _fields(mm)
    assert float(np.max(np.abs(E_r - E_f))) < 1e-12 * float(np.max(np.abs(E_f)))
    assert float(np.max(np.abs(P_r - P_f))) < 1e-12 * float(np.max(np.abs(P_f)))


def test_field_changes_keep_the_shadow():
    """A Zeeman field change (as a clocking protocol makes every field step) does not enter the
    dipolar correlations: the next periodic recompute still refreshes, with the new field."""
    mm, step = _build("neel_pinwheel_perp", seed=4)
    z = mm.get_energy('zeeman')
    for k in range(3):
        z.set_field(1e-3 * (1 + k), 0.3 * k)
        _step_some(mm, step)
        mm._device_recompute(periodic=True)
        assert mm._dev.shadow_refreshes == 1  # each full recompute below stores a new shadow
        E_r, P_r = _fields(mm)
        mm._device_recompute()
        E_f, P_f = _fields(mm)
        assert float(np.max(np.abs(E_r - E_f))) < 1e-12 * float(np.max(np.abs(E_f)))
        assert float(np.max(np.abs(P_r - P_f))) < 1e-12 * float(np.max(np.abs(P_f)))


def test_periodic_recomputes_refresh_in_progress():
    """progress() with recompute_every: the periodic recomputes refresh from the shadow, at most
    SHADOW_MAX_REFRESHES in a row, and the fields stay those of a full recompute."""
    mm, _ = _build("neel_pinwheel_perp", seed=5)
    mm.recompute_every = 32
    mm._dev.SHADOW_MAX_REFRESHES = 4
    mm.progress(t_max=math.inf, MCsteps_max=0.25)
    assert mm.recomputes >= 6
    assert mm._dev.refreshes >= 4 and mm._dev.shadow_refreshes <= 4
    E_i, P_i = _fields(mm)
    mm._device_recompute()
    E_f, P_f = _fields(mm)
    assert float(np.max(np.abs(E_i - E_f))) < 1e-5 * float(np.max(np.abs(E_f)))
    assert float(np.max(np.abs(P_i - P_f))) < 1e-5 * float(np.max(np.abs(P_f)))


def _pin64(seed, every):
    import hotspice_amd as hs
    n = 64
    return hs.ASI.IP_Pinwheel(1e-6, n, T=300, moment=2.37e-16, seed=seed, recompute_every=every,
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(1e-3, math.pi / 8)],
                              params=hs.SimParams(UPDATE_SCHEME='Néel'), **_inputs((n, n), seed, EB_kT=40))


def _dump(mm):
    d = mm._dev
    return [t.cpu().numpy().copy() for t in (d.m, d.E, d.Ec, d.Ep, d.Epc)] + [mm.t, mm.switches]


def test_batched_refreshes_equal_alone():
    """Neel instances stepped in one batched launch, with their periodic recomputes queued behind
    it (batch.progress_many), take the same refresh-or-full decisions as each stepped alone
    through progress(), so both end bit-identical; refreshes do happen."""
    from hotspice_amd import batch
    seeds = [21, 22, 23]
    alone = [_pin64(s, 96) for s in seeds]
    together = [_pin64(s, 96) for s in seeds]
    for _ in range(4):
        for m in alone:
            m.progress(t_max=math.inf, MCsteps_max=300 / m.n)
        batch.progress_many(together, t_max=math.inf, MCsteps_max=300 / together[0].n)
    for a, b in zip(alone, together):
        assert a._dev.refreshes > 0 and a._dev.refreshes == b._dev.refreshes
        assert a.recomputes == b.recomputes
        for x, y in zip(_dump(a), _dump(b)):
            if isinstance(x, np.ndarray):
                np.testing.assert_array_equal(x, y)
            else:
                assert x == y


def test_refresh_entry_point_refusals():
    """hs_refresh_fields refuses an instance without a shadow and a `which` that leaves out a
    component the shadow holds (E_perp of an in-plane instance)."""
    import ctypes as C
    from hotspice_amd import _lib
    mm, _ = _build("neel_pinwheel_perp", seed=9)
    d = mm._dev
    scr = d.recompute_scratch()
    with pytest.raises(_lib.HotspiceError, match="covers every field component"):
        _lib.check(d.lib.hs_refresh_fields(C.byref(d.inst), 1, d.ptr(scr), scr.numel(), d.stream()))
    with pytest.raises(_lib.HotspiceError, match="scratch"):
        _lib.check(d.lib.hs_refresh_fields(C.byref(d.inst), 3, d.ptr(scr), 16, d.stream()))
    inst = _lib.Instance.from_buffer_copy(bytes(d.inst))
    inst.C64 = None
    with pytest.raises(_lib.HotspiceError, match="no FP64 shadow"):
        _lib.check(d.lib.hs_refresh_fields(C.byref(inst), 3, d.ptr(scr), scr.numel(), d.stream()))
    assert int(d.lib.hs_refresh_scratch_bytes(C.byref(d.inst))) == 12 * mm.nx * mm.ny + 8 * mm.ny + 4
