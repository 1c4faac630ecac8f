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

import numpy as np

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
    if scheme == Scheme.NEEL and not all(m._uniform_beta for m in mags):
        # a temperature that varies over the grid steps through the per-step launches (the
        # persistent team kernel needs a uniform one): those instances go one at a time
        nbl = nblocks if isinstance(nblocks, (list, tuple)) else [nblocks] * len(mags)
        out = [None] * len(mags)
        uni = [i for i, m in enumerate(mags) if m._uniform_beta]
        if uni:
            res = progress_many([mags[i] for i in uni], t_max, MCsteps_max, [nbl[i] for i in uni], chunk,
                                None if per_instance is None else [per_instance[i] for i in uni], **kwargs)
            for i, r in zip(uni, res):
                out[i] = r
        for i, m in enumerate(mags):
            if out[i] is None:
                kw = dict(kwargs, **(per_instance[i] if per_instance is not None else {}))
                out[i] = m.progress(t_max=t_max, MCsteps_max=MCsteps_max, **kw)
        return out
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
    sts = read_states(mags)
    starts = [(st.t, st.attempts) for st in sts]
    attempt_freq = float(kwargs.get('attempt_freq', 1e10))
    if per_instance is not None and len(per_instance) != len(mags):
        raise ValueError("per_instance needs one keyword dict per instance")
    kws = [dict(kwargs, **(per_instance[i] if per_instance is not None else {})) for i in range(len(mags))]
    if scheme == Scheme.NEEL:
        return _progress_many_neel(mags, starts, t_max, MCsteps_max, attempt_freq, nbl, chunk, sts)
    if scheme == Scheme.METROPOLIS:
        return _progress_many_metropolis(mags, starts, t_max, MCsteps_max, attempt_freq, chunk, kws, sts)
    raise NotImplementedError("batched Wolff updates are not supported")


def read_states(mags):
    """The device run states of many instances with one gather and one device-to-host copy."""
    if not mags:
        return []
    torch = mags[0]._dev.torch
    raw = torch.stack([m._dev.state for m in mags]).cpu().numpy()[:, :_lib.STATE_BYTES]
    return [_lib.State.from_buffer_copy(row) for row in np.ascontiguousarray(raw)]


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


def launch_neel(items, attempt_freq=1e10, begin=False, full=None, updates=False):
    """ONE persistent launch stepping every (mm, max_steps, t_max, mc_max, nblocks) item. With
    begin=True the launch starts each instance's progress call (its origin t0, attempts0 := the
    current t, attempts); otherwise the origin must have been set. updates=True runs max_steps
    update(t_max) calls instead of a progress() call: every step runs, whatever t has reached
    (also +inf), and compares its waiting time with t_max itself (core.py:702). Afterwards checks
    each instance's error word (raising on the first), counts its steps towards the periodic FP64
    recompute and runs that when due. `full[i]`: item i is known to run all its max_steps (no
    stop predicate can end it earlier); when those steps complete its recompute period, the
    recompute is queued right behind the launch instead of after reading the states back (the
    same recompute at the same point of the trajectory)."""
    if not items:
        return
    queued = enqueue_neel(items, attempt_freq, begin, free=[2 if updates else 0] * len(items), full=full)
    sts, errs = collect([it[0] for it in items], recomputed={id(m) for m in queued})
    _raise_first(errs, "Neel")
    return sts


def enqueue_neel(items, attempt_freq=1e10, begin=False, free=None, full=None):
    """The launches of launch_neel on the current stream, without reading anything back.
    free[i]: hs_run_t.first_free of item i (0: a progress() call, 2: update() calls). begin: one
    flag for every item or one per item. Returns the instances whose recompute was queued behind
    the launch (see launch_neel's `full`)."""
    mags = [it[0] for it in items]
    dev = mags[0]._dev
    for m in mags:
        if not m._uniform_beta:
            raise NotImplementedError("the batched Neel launch (persistent team kernel) needs a uniform temperature; "
                                      "progress_many steps instances with a varying one through Magnets.progress")
    free = [0] * len(items) if free is None else list(free)
    begin = [int(begin)] * len(items) if np.ndim(begin) == 0 else [int(b) for b in begin]
    # one launch per E_perp setting: the team kernel is compiled for either, and instances are
    # independent, so two launches in stream order step each exactly as one would
    perp = [bool(m.USE_PERP_ENERGY) for m in mags]
    for flag in sorted(set(perp)):
        g = [i for i, p in enumerate(perp) if p == flag]
        teams = _team_array([mags[i] for i in g], [items[i][1] for i in g], [items[i][2] for i in g],
                            [items[i][3] for i in g], attempt_freq, [begin[i] for i in g], nblocks=[items[i][4] for i in g],
                            first_free=[free[i] for i in g])
        dev.run_teams(teams, "neel")
    queued = []
    if full is not None:
        due = [(m, steps) for (m, steps, *_), f in zip(items, full)
               if f and m.recompute_every and m._since_recompute + steps >= m.recompute_every]
        queued = [m for m, _ in due]
        if queued:
            dev.recompute_many([m._dev for m in queued], [3 if m.USE_PERP_ENERGY else 1 for m in queued],
                               [m._changed_bound(ahead=steps) for m, steps in due])
    return queued


_TEAM_DTYPE = np.dtype(_lib.Team)


def _team_array(mags, steps, t_max, mc, attempt_freq, begin, nblocks=None, sels=None, first_free=0, grid=None):
    """The hs_team_t descriptors of one launch, filled column by column (`nblocks`: Neel team
    sizes; `grid`: Metropolis update grids, hs_run_t.nblocks)."""
    arr = np.zeros(len(mags), dtype=_TEAM_DTYPE)
    arr['inst'] = [_lib.C.addressof(m._dev.inst) for m in mags]
    arr['state'] = [m._dev.state.data_ptr() for m in mags]
    if sels is not None:
        arr['sel'] = [_lib.C.addressof(s) for s in sels]
    run = arr['run']
    run['t_max'] = np.asarray(t_max, dtype=float)
    run['mc_max'] = np.asarray(mc, dtype=float)
    run['attempt_freq'] = float(attempt_freq)
    run['max_steps'] = np.asarray(steps, dtype=np.int64)
    run['begin'] = np.asarray(begin, dtype=np.int32)  # one flag for all, or one per instance
    run['first_free'] = np.asarray(first_free, dtype=np.int32)
    if nblocks is not None:
        arr['nblocks'] = np.asarray(nblocks, dtype=np.int32)
    if grid is not None:
        run['nblocks'] = np.asarray(grid, dtype=np.int32)
    return arr


def launch_metropolis(items, attempt_freq=1e10, begin=False):
    """One batched Metropolis call stepping every (mm, sel, max_steps, t_max, mc_max) item: a
    selection launch and an update launch per step for all instances together (begin: as for
    launch_neel)."""
    if not items:
        return
    mags = [it[0] for it in items]
    before = [st.attempts for st in read_states(mags)]
    enqueue_metropolis(items, attempt_freq, begin)
    sts, errs = collect(mags, before)
    _raise_first(errs, "Metropolis")
    return sts


def enqueue_metropolis(items, attempt_freq=1e10, begin=False):
    """The launches of launch_metropolis on the current stream, without reading anything back
    (begin: one flag for every item or one per item)."""
    mags = [it[0] for it in items]
    teams = _team_array(mags, [it[2] for it in items], [it[3] for it in items], [it[4] for it in items],
                        attempt_freq, begin, sels=[it[1] for it in items], grid=[m._metro_grid() for m in mags])
    mags[0]._dev.run_teams(teams, "metropolis")


def _raise_first(errs, what):
    for i, e in errs.items():
        raise _lib.HotspiceError(f"device {what} step failed (err={e})")


def collect(mags, before=None, recomputed=()):
    """After a launch: reads the states back, records samples per step (`before`: attempts before
    the launch) and runs the recomputes that are due (except those already queued behind the
    launch, `recomputed`: ids of instances). Returns (states, {index: device error word}) -- an
    instance with an error is left as it is (no recompute), so its caller can drop it."""
    sts = read_states(mags)
    due, errs = [], {}
    for i, (m, st) in enumerate(zip(mags, sts)):
        if st.err:
            errs[i] = int(st.err)
            continue
        if before is not None and st.steps_done > 0:  # samples per step, to size the next launch
            m._metro_rate = max(1.0, (st.attempts - before[i]) / st.steps_done)
        m._since_recompute += int(st.steps_done)
        if id(m) in recomputed:
            m._recompute_done(st)
            continue
        if m.recompute_every and m._since_recompute >= m.recompute_every and m._incremental_exact():
            due.append(m)
    if due:  # the periodic FP64 recomputes of all instances due, in one call (or refreshes)
        stm = {id(m): st for m, st in zip(mags, sts)}
        due[0]._dev.recompute_many([m._dev for m in due], [3 if m.USE_PERP_ENERGY else 1 for m in due],
                                   [m._changed_bound(stm[id(m)]) for m in due])
        for m in due:
            m._recompute_done(stm[id(m)])
    return sts, errs


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


def _lower_than_vec(x, y, rtol=1e-10):
    """lower_than (utils.py:90-92) over arrays."""
    x, y = np.asarray(x, dtype=float), np.asarray(y, dtype=float)
    return (x < y) & ~np.isclose(x, y, rtol=rtol, atol=0)


def _progress_many_neel(mags, starts, t_max, mc, f0, nbl, chunk, sts):
    """The pending test and the chunk sizes of neel_chunk, over all instances at once. A launch in
    which every instance runs all its steps (t_max infinite, MC budget lasting past them) is
    queued without reading the states back: each step is one attempt, so the host advances the
    attempt counts itself (t does not enter the pending test then), and the states are read once
    a launch may stop early or the call ends."""
    first = True
    n = np.array([m.n for m in mags], dtype=float)
    t0 = np.array([s[0] for s in starts], dtype=float)
    a0 = np.array([s[1] for s in starts], dtype=np.int64)
    every = np.array([m.recompute_every or 0 for m in mags], dtype=np.int64)
    t = np.array([st.t for st in sts], dtype=float)
    att = np.array([st.attempts for st in sts], dtype=np.int64)
    unread = False  # launches queued since the states were last read
    while True:
        pend = _lower_than_vec(t - t0, t_max) & _lower_than_vec(att / n - a0 / n, mc)
        idx = np.flatnonzero(pend)
        if idx.size == 0:
            break
        if math.isfinite(mc):
            steps = np.ceil(mc * n[idx] - (att[idx] - a0[idx])).astype(np.int64) + 1
        else:
            steps = np.full(idx.size, 4096, dtype=np.int64)
        if chunk:
            steps = np.minimum(steps, chunk)
        if math.isfinite(t_max):
            steps = np.minimum(steps, 4096)
        since = np.array([mags[i]._since_recompute for i in idx], dtype=np.int64)
        ev = every[idx]
        steps = np.where(ev > 0, np.minimum(steps, np.maximum(1, ev - since)), steps)
        steps = np.maximum(1, steps)
        items = [(mags[i], int(k), t_max, mc, nbl[i]) for i, k in zip(idx.tolist(), steps.tolist())]
        # an instance runs all its steps when t_max is infinite and its MC budget lasts past them
        if math.isinf(t_max) and t_max > 0:
            left = np.ceil(mc * n[idx] - (att[idx] - a0[idx])) if math.isfinite(mc) else np.full(idx.size, np.inf)
            full = (steps <= left).tolist()
        else:
            full = None
        # the first launch opens every pending instance's progress call (origin = starts)
        if full is not None and all(full):
            queued = {id(m) for m in enqueue_neel(items, f0, begin=first, full=full)}
            first = False
            unread = True
            for i, k in zip(idx.tolist(), steps.tolist()):
                att[i] += k
                m = mags[i]
                if id(m) in queued:
                    m._recompute_done()
                else:
                    m._since_recompute += k
            continue
        after = launch_neel(items, f0, begin=first, full=full)
        first = False
        for i, st in zip(idx.tolist(), after):
            sts[i] = st
        if unread and idx.size < len(mags):  # the others advanced in the queued launches
            in_launch = set(idx.tolist())
            fresh = read_states(mags)
            _raise_first({i: int(st.err) for i, st in enumerate(fresh) if st.err}, "Neel")
            for i in range(len(mags)):
                if i not in in_launch:
                    sts[i] = fresh[i]
        unread = False
        t = np.array([st.t for st in sts], dtype=float)
        att = np.array([st.attempts for st in sts], dtype=np.int64)
    if unread:
        sts = read_states(mags)
        _raise_first({i: int(st.err) for i, st in enumerate(sts) if st.err}, "Neel")
    return _results(mags, starts, sts)


def _progress_many_metropolis(mags, starts, t_max, mc, f0, chunk, kws, sts):
    sels = []
    for m, kw in zip(mags, kws):
        if kw.get('idx') is not None:  # the instance's own sample list, every step
            sel = m._explicit_selection(kw['idx'])
            if sel is None:
                raise ValueError("progress() with an empty idx never advances t or MCsteps")
            sels.append(sel)
            continue
        r = kw.get('r', 0)
        r = m.calc_r(kw.get('Q', 0.05)) if r == 0 else r
        m._metro_nblocks(r)  # 'auto' instances are tuned first
        sels.append(m._select_geometry(r))
    first = True
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
