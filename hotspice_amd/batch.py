"""Batched stepping: many independent Magnets instances in one persistent launch.

progress_many([...]) runs Magnets.progress() (hotspice/core.py:736-760) on every instance,
with each instance's own t_max / MCsteps_max, temperature, fields and Philox seed, in ONE
cooperative launch per chunk: every instance is a team of workgroups that steps on its own,
so the results are those of stepping each instance alone (the Philox streams are keyed by
the instance's seed and step counter, not by its position in the batch).
"""
import math

from . import _lib
from .core import Scheme
from .utils import lower_than


def _check_same_scheme(mags):
    schemes = {m.params.UPDATE_SCHEME for m in mags}
    if len(schemes) != 1:
        raise ValueError("progress_many needs instances of one UPDATE_SCHEME")
    return schemes.pop()


def progress_many(magnets, t_max=1.0, MCsteps_max=4.0, nblocks=None, chunk=None, per_instance=None, **kwargs):
    """Steps every instance like mm.progress(t_max, MCsteps_max, **kwargs), all instances
    together per chunk. `nblocks`: Neel workgroups per instance (int, list, or None for
    automatic). `per_instance`: optional list of keyword dicts (e.g. Metropolis r / Q) that
    override **kwargs for one instance each. Returns the list of (dt, dMCsteps)."""
    mags = list(magnets)
    if not mags:
        return []
    scheme = _check_same_scheme(mags)
    dev = mags[0]._dev
    for m in mags:
        if m._dev.device != dev.device:
            raise ValueError("all instances of a batch must live on one device")
        if m._T_any_zero:
            raise ValueError("T=0 somewhere in an instance: no switch can be simulated")
        m._pre_step()
    nbl = nblocks if isinstance(nblocks, (list, tuple)) else [nblocks or 0] * len(mags)
    starts = []
    for m in mags:
        st = m._state()
        starts.append((st.t, st.attempts))
    attempt_freq = float(kwargs.get('attempt_freq', 1e10))
    if per_instance is not None and len(per_instance) != len(mags):
        raise ValueError("per_instance needs one keyword dict per instance")
    kws = [dict(kwargs, **(per_instance[i] if per_instance is not None else {})) for i in range(len(mags))]
    if scheme == Scheme.NEEL:
        return _progress_many_neel(mags, starts, t_max, MCsteps_max, attempt_freq, nbl, chunk)
    if scheme == Scheme.METROPOLIS:
        return _progress_many_metropolis(mags, starts, t_max, MCsteps_max, attempt_freq, chunk, kws)
    raise NotImplementedError("batched Wolff updates are not supported")


def _pending(m, start, t_max, mc):
    st = m._state()
    t0, a0 = start
    return lower_than(st.t - t0, t_max) and lower_than(st.attempts / m.n - a0 / m.n, mc), st


def _progress_many_neel(mags, starts, t_max, mc, f0, nbl, chunk):
    dev = mags[0]._dev
    first = True
    while True:
        teams = []
        for m, s0, nb in zip(mags, starts, nbl):
            pend, st = _pending(m, s0, t_max, mc)
            if not pend:
                continue
            if not m._uniform_beta:
                raise NotImplementedError("Neel stepping on the device needs a spatially uniform temperature.")
            if math.isfinite(mc):
                steps = int(math.ceil(mc * m.n - (st.attempts - s0[1]))) + 1
            else:
                steps = 4096
            if chunk:
                steps = min(steps, chunk)
            if math.isfinite(t_max):
                steps = min(steps, 4096)
            if m.recompute_every:
                steps = min(steps, max(1, m.recompute_every - m._since_recompute))
            if first:
                m._set_progress_origin(*s0)
            run = _lib.Run(t_max=float(t_max), mc_max=float(mc), attempt_freq=f0, max_steps=steps, log=0, nblocks=0,
                           begin=0)
            teams.append((m, _lib.Team(inst=_lib.C.addressof(m._dev.inst), sel=None, run=run,
                                       state=m._dev.state.data_ptr(), log_idx=None, log_dt=None, log_cells=None,
                                       nblocks=int(nb))))
        first = False
        if not teams:
            break
        dev.run_teams([t for _, t in teams], "neel")
        for m, _ in teams:
            st = m._state()
            if st.err:
                raise _lib.HotspiceError(f"device Neel step failed (err={st.err})")
            m._since_recompute += int(st.steps_done)
            m._maybe_recompute()
    out = []
    for m, (t0, a0) in zip(mags, starts):
        m._after_device_step()
        st = m._state()
        out.append((st.t - t0, st.attempts / m.n - a0 / m.n))
    return out


def _progress_many_metropolis(mags, starts, t_max, mc, f0, chunk, kws):
    dev = mags[0]._dev
    sels = []
    for m, kw in zip(mags, kws):
        if kw.get('idx') is not None:
            raise NotImplementedError("Explicit sample lists (idx=...) are not supported by the device Metropolis step.")
        r = kw.get('r', 0)
        Q = kw.get('Q', 0.05)
        sels.append(m._select_geometry(m.calc_r(Q) if r == 0 else r))
    first = True
    while True:
        teams = []
        for m, s0, sel in zip(mags, starts, sels):
            pend, st = _pending(m, s0, t_max, mc)
            if not pend:
                continue
            per_step = max(1.0, 0.25 * sel.nsc)
            if math.isfinite(mc):
                steps = int(math.ceil((mc * m.n - (st.attempts - s0[1])) / per_step)) + 1
                steps = max(1, min(steps, 1 << 16))
            else:
                steps = 1024
            if chunk:
                steps = min(steps, chunk)
            if m.recompute_every and m._incremental_exact():
                steps = min(steps, max(1, m.recompute_every - m._since_recompute))
            if first:
                m._set_progress_origin(*s0)
            run = _lib.Run(t_max=float(t_max), mc_max=float(mc), attempt_freq=f0, max_steps=steps, log=0, nblocks=0,
                           begin=0)
            teams.append((m, _lib.Team(inst=_lib.C.addressof(m._dev.inst), sel=_lib.C.addressof(sel), run=run,
                                       state=m._dev.state.data_ptr(), log_idx=None, log_dt=None, log_cells=None,
                                       nblocks=0)))
        first = False
        if not teams:
            break
        dev.run_teams([t for _, t in teams], "metropolis")
        for m, _ in teams:
            st = m._state()
            if st.err:
                raise _lib.HotspiceError(f"device Metropolis step failed (err={st.err})")
            m._since_recompute += int(st.steps_done)
            m._maybe_recompute()
    out = []
    for m, (t0, a0) in zip(mags, starts):
        m._after_device_step()
        st = m._state()
        out.append((st.t - t0, st.attempts / m.n - a0 / m.n))
    return out
