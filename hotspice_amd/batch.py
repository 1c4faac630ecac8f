"""Batched stepping: many independent Magnets instances advanced together.

progress_many([...]) runs Magnets.progress() (hotspice/core.py:736-760) on every instance,
with each instance's own t_max / MCsteps_max, temperature, fields and Philox seed. Neel
instances step in ONE persistent cooperative launch per chunk (every instance is a team of
workgroups that steps on its own); Metropolis instances step together, one selection launch
and one update launch per step for all of them. Results are those of stepping each instance
alone: the Philox streams are keyed by the instance's seed and step counter, not by its
position in the batch, and every instance's launches are chunked as Magnets.progress chunks
them. launch_neel / launch_metropolis are the single-launch building blocks (used by serve.py).
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
    if scheme == Scheme.NEEL:  # per-instance team sizes ('auto' instances are tuned first)
        nbl = [nb or (m._neel_nblocks() if m.neel_blocks else 0) for m, nb in zip(mags, nbl)]
    starts = [(st.t, st.attempts) for st in read_states(mags)]
    attempt_freq = float(kwargs.get('attempt_freq', 1e10))
    if per_instance is not None and len(per_instance) != len(mags):
        raise ValueError("per_instance needs one keyword dict per instance")
    kws = [dict(kwargs, **(per_instance[i] if per_instance is not None else {})) for i in range(len(mags))]
    if scheme == Scheme.NEEL:
        return _progress_many_neel(mags, starts, t_max, MCsteps_max, attempt_freq, nbl, chunk)
    if scheme == Scheme.METROPOLIS:
        return _progress_many_metropolis(mags, starts, t_max, MCsteps_max, attempt_freq, chunk, kws)
    raise NotImplementedError("batched Wolff updates are not supported")


def read_states(mags):
    """The device run states of many instances with one gather and one device-to-host copy."""
    if not mags:
        return []
    torch = mags[0]._dev.torch
    raw = torch.stack([m._dev.state[:_lib.STATE_BYTES] for m in mags]).cpu().numpy()
    return [_lib.State.from_buffer_copy(raw[i].tobytes()) for i in range(len(mags))]


def _pending(m, start, t_max, mc, st=None):
    st = m._state() if st is None else st
    t0, a0 = start
    return lower_than(st.t - t0, t_max) and lower_than(st.attempts / m.n - a0 / m.n, mc), st


def neel_team_size(m, nblocks=0):
    """Workgroups of an instance's Neel team: `nblocks` (or the instance's own setting), else the
    library's default of one magnet per owner thread (512 per workgroup), at most one per CU."""
    nblocks = nblocks or (m.neel_blocks if isinstance(m.neel_blocks, int) else 0)
    if nblocks:
        return int(nblocks)
    cus = m._dev.torch.cuda.get_device_properties(m._dev.device).multi_processor_count
    return max(1, min(cus, -(-m.n // 512)))


def launch_neel(items, attempt_freq=1e10, begin=False):
    """ONE persistent launch stepping every (mm, max_steps, t_max, mc_max, nblocks) item. With
    begin=True the launch starts each instance's progress call (its origin t0, attempts0 := the
    current t, attempts); otherwise the origin must have been set. Afterwards checks each
    instance's error word, counts its steps towards the periodic FP64 recompute and runs that
    when due."""
    if not items:
        return
    dev = items[0][0]._dev
    teams = []
    for m, steps, t_max, mc, nb in items:
        if not m._uniform_beta:
            raise NotImplementedError("Neel stepping on the device needs a spatially uniform temperature.")
        run = _lib.Run(t_max=float(t_max), mc_max=float(mc), attempt_freq=float(attempt_freq), max_steps=int(steps),
                       log=0, nblocks=0, begin=int(begin))
        teams.append(_lib.Team(inst=_lib.C.addressof(m._dev.inst), sel=None, run=run, state=m._dev.state.data_ptr(),
                               log_idx=None, log_dt=None, log_cells=None, nblocks=int(nb)))
    dev.run_teams(teams, "neel")
    return _after_launch([it[0] for it in items], "Neel")


def launch_metropolis(items, attempt_freq=1e10, begin=False):
    """One batched Metropolis call stepping every (mm, sel, max_steps, t_max, mc_max) item: a
    selection launch and an update launch per step for all instances together (begin: as for
    launch_neel)."""
    if not items:
        return
    dev = items[0][0]._dev
    teams = []
    for m, sel, steps, t_max, mc in items:
        run = _lib.Run(t_max=float(t_max), mc_max=float(mc), attempt_freq=float(attempt_freq), max_steps=int(steps),
                       log=0, nblocks=0, begin=int(begin))
        teams.append(_lib.Team(inst=_lib.C.addressof(m._dev.inst), sel=_lib.C.addressof(sel), run=run,
                               state=m._dev.state.data_ptr(), log_idx=None, log_dt=None, log_cells=None, nblocks=0))
    before = [st.attempts for st in read_states([it[0] for it in items])]
    dev.run_teams(teams, "metropolis")
    return _after_launch([it[0] for it in items], "Metropolis", before)


def _after_launch(mags, what, before=None):
    """Checks errors, records samples per step and runs due recomputes; returns the states."""
    sts = read_states(mags)
    due = []
    for i, (m, st) in enumerate(zip(mags, sts)):
        if st.err:
            raise _lib.HotspiceError(f"device {what} step failed (err={st.err})")
        if before is not None and st.steps_done > 0:  # samples per step, to size the next launch
            m._metro_rate = max(1.0, (st.attempts - before[i]) / st.steps_done)
        m._since_recompute += int(st.steps_done)
        if m.recompute_every and m._since_recompute >= m.recompute_every and m._incremental_exact():
            due.append(m)
    if due:  # the periodic FP64 recomputes of all instances due, in one call
        due[0]._dev.recompute_many([m._dev for m in due], [3 if m.USE_PERP_ENERGY else 1 for m in due])
        for m in due:
            m._since_recompute = 0
    return sts


def neel_chunk(m, start, t_max, mc, chunk=None, st=None):
    """Steps of the next Neel launch for an instance inside progress(t_max, mc) started at
    `start` = (t0, attempts0): as Magnets.progress chunks it."""
    st = m._state() if st is None else st
    if math.isfinite(mc):
        steps = int(math.ceil(mc * m.n - (st.attempts - start[1]))) + 1
    else:
        steps = 4096
    if chunk:
        steps = min(steps, chunk)
    if math.isfinite(t_max):
        steps = min(steps, 4096)
    if m.recompute_every:
        steps = min(steps, max(1, m.recompute_every - m._since_recompute))
    return max(1, steps)


def metropolis_chunk(m, sel, start, mc, chunk=None, st=None):
    """Steps of the next Metropolis launch (as Magnets.progress chunks it)."""
    st = m._state() if st is None else st
    per_step = m._metro_rate or max(1.0, 0.25 * sel.nsc)
    if math.isfinite(mc):
        steps = int(math.ceil((mc * m.n - (st.attempts - start[1])) / per_step)) + 1
        steps = max(1, min(steps, 1 << 16))
    else:
        steps = 1024
    if chunk:
        steps = min(steps, chunk)
    if m.recompute_every and m._incremental_exact():
        steps = min(steps, max(1, m.recompute_every - m._since_recompute))
    return max(1, steps)


def _progress_many_neel(mags, starts, t_max, mc, f0, nbl, chunk):
    first = True
    sts = read_states(mags)
    while True:
        items = [(m, neel_chunk(m, s0, t_max, mc, chunk, st), t_max, mc, nb)
                 for m, s0, nb, st in zip(mags, starts, nbl, sts) if _pending(m, s0, t_max, mc, st)[0]]
        if not items:
            break
        # the first launch opens every pending instance's progress call (origin = starts)
        after = dict(zip((id(it[0]) for it in items), launch_neel(items, f0, begin=first)))
        first = False
        sts = [after.get(id(m), st) for m, st in zip(mags, sts)]
    return _results(mags, starts, sts)


def _progress_many_metropolis(mags, starts, t_max, mc, f0, chunk, kws):
    sels = []
    for m, kw in zip(mags, kws):
        if kw.get('idx') is not None:
            raise NotImplementedError("Explicit sample lists (idx=...) are not supported by the device Metropolis step.")
        r = kw.get('r', 0)
        Q = kw.get('Q', 0.05)
        sels.append(m._select_geometry(m.calc_r(Q) if r == 0 else r))
    first = True
    sts = read_states(mags)
    while True:
        items = [(m, sel, metropolis_chunk(m, sel, s0, mc, chunk, st), t_max, mc)
                 for m, s0, sel, st in zip(mags, starts, sels, sts) if _pending(m, s0, t_max, mc, st)[0]]
        if not items:
            break
        after = dict(zip((id(it[0]) for it in items), launch_metropolis(items, f0, begin=first)))
        first = False
        sts = [after.get(id(m), st) for m, st in zip(mags, sts)]
    return _results(mags, starts, sts)


def _results(mags, starts, sts=None):
    out = []
    sts = read_states(mags) if sts is None else sts
    for m, (t0, a0), st in zip(mags, starts, sts):
        m._after_device_step()
        out.append((st.t - t0, st.attempts / m.n - a0 / m.n))
    return out
