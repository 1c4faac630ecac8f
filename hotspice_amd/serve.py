"""Continuous-batching server for field-step jobs on one device.

A job is a spin-ice instance driven through a field program, the pattern of the reference's
inputters (FieldInputter.input_single, hotspice/io.py:221-239; the clocking examples,
examples/Clocking_IP_Pinwheel.py:28-55): a list of io.FieldStep records, each a
ZeemanEnergy.set_field call (energies.py:101-114) followed by `count` update(t_max) calls
(Neel, core.py:679-709) or a progress(t_max, MCsteps_max) call (core.py:736-760). Any inputter's
program can be served (Inputter.schedule); the short form Job(fields=[(magnitude, angle)...],
work=w) means `w` update(t_max=inf) calls per field step under Neel and progress(t_max=inf,
MCsteps_max=w) under Metropolis.

The server
  * admits queued jobs whose arrival time has passed while their HBM footprint fits the budget
    (state arrays + dipolar kernel tables, estimated before the instance is built);
  * keeps every admitted job's state resident on the device between iterations;
  * advances all active jobs together in each iteration: one persistent Neel launch (teams
    packed up to the co-resident workgroup limit, the rest wait for the next iteration) and one
    batched Metropolis call, each job running at most to the end of its current field step; with
    overlap=True the two go on two streams of the default priority, so their kernels can share
    the device;
  * admits jobs that arrived meanwhile at the next iteration (continuous batching), frees a job's
    device memory when its last field step is done and records its latency;
  * isolates failures: a job whose instance cannot be built, whose program cannot be served, whose
    descriptor the library rejects or whose device step reports an error is marked failed with
    the error, its memory is freed and the other jobs go on -- the reference's job runner records
    a failed job and carries on the same way (ParallelJobs.py:93-95, 108-112).

Every launch of a job is chunked from the job's own state only (its step's work, its recompute
period), and the kernels give an instance the same arithmetic alone or in a batch, so a job's
result does not depend on what else is being served.
"""
import dataclasses
import heapq
import math
import time
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import _lib, batch
from .core import Scheme
from .io import FieldStep
from .utils import lower_than


@dataclasses.dataclass
class Job:
    """One request. `make()` builds its Magnets (on the device). The work is `program` (a list of
    io.FieldStep), or, when that is None, the field steps `fields` = [(magnitude, angle), ...]
    with `work` Neel update() calls or Metropolis MC sweeps each."""
    make: Callable
    fields: List[Tuple[float, float]] = dataclasses.field(default_factory=list)
    work: float = 0.0
    arrival: float = 0.0
    footprint: int = 0
    name: str = ""
    kwargs: dict = dataclasses.field(default_factory=dict)  # Metropolis r / Q, attempt_freq
    spec: Optional[dict] = None  # what make() builds (for the caller's records; unused here)
    program: Optional[list] = None
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[dict] = None
    status: str = "queued"  # queued -> running -> done | failed
    error: Optional[str] = None


def footprint_bytes(nx, ny, n_sub, kernel_full=True, kwx=None, kwy=None):
    """Estimated device bytes of one instance: per cell int8 m + uint8 flags + FP32 E, Ec, Ep, Epc,
    E_B, beta, mu + FP64 mu (38 B), per sublattice the FP32 K, K_perpother and FP64 K, K_perpself
    tables over the stored window (24 B per entry), and the FP64 recompute scratch (two FP64
    planes per source group)."""
    if kernel_full or kwx is None:
        kwx, kwy = nx - 1, ny - 1
    ncell = nx * ny
    split = min(64, max(1, math.ceil(768 * 2048 / ncell)))  # source-group split of the recompute
    return ncell * 38 + n_sub * (2 * kwx + 1) * (2 * kwy + 1) * 24 + ncell * 16 * split


def job_program(job: Job, scheme) -> list:
    """The field steps of a job (its own program, or the one its fields / work stand for)."""
    if job.program is not None:
        return list(job.program)
    if scheme == Scheme.NEEL:
        return [FieldStep(magnitude=mag, angle=ang, advance='update', count=int(job.work), t_max=math.inf)
                for mag, ang in job.fields]
    return [FieldStep(magnitude=mag, angle=ang, advance='progress', t_max=math.inf, MCsteps_max=float(job.work))
            for mag, ang in job.fields]


class JobError(Exception):
    """A job that cannot be served (raised inside the server and recorded on the job)."""


class _Active:
    """A running job: its instance, its field steps and where it is in them."""

    def __init__(self, job, mm, now, backend=batch):
        self.job = job
        self.mm = mm
        self.batch = backend
        self.scheme = mm.params.UPDATE_SCHEME
        self.steps = job_program(job, self.scheme)
        for st in self.steps:
            if st.advance == 'minimize':
                raise JobError("minimize() steps cannot be served (a zero-temperature sweep runs on its own)")
            if st.advance == 'update' and self.scheme != Scheme.NEEL:
                raise JobError("update() steps are served for Neel jobs only; give Metropolis steps as progress()")
        if self.scheme not in (Scheme.NEEL, Scheme.METROPOLIS):
            raise JobError(f"{self.scheme} jobs cannot be served")
        if self.scheme == Scheme.NEEL and not mm._uniform_beta:
            raise JobError("served Neel jobs run in the persistent team kernel, which needs a uniform temperature")
        self.sel = None
        if self.scheme == Scheme.METROPOLIS:
            r = job.kwargs.get('r', 0)
            Q = job.kwargs.get('Q', 0.05)
            self.sel = mm._select_geometry(mm.calc_r(Q) if r == 0 else r)
        self.zeeman = mm.get_energy('Zeeman', verbose=False)
        if self.zeeman is None:
            raise JobError("a served job's instance needs a ZeemanEnergy")
        self.index = 0
        # the field step's first launch sets its progress origin on the device (hs_run_t.begin:
        # t0, attempts0 := t, attempts); until it has run, the step cannot be done
        self.pending_begin = False
        self.st = None  # the job's state as last read back (by the admission or a collect)
        job.admitted = now
        job.status = "running"

    @property
    def current(self) -> FieldStep:
        return self.steps[self.index]

    def begin(self):
        """Applies field steps from the current one on: sets the field, and stops at the first step
        that advances the instance, whose first launch will record its origin (t, attempts) on the
        device. Returns True when the program is complete."""
        while self.index < len(self.steps):
            st = self.current
            self.zeeman.set_field(magnitude=st.magnitude, angle=st.angle)
            if st.advance != 'none':
                self.mm._pre_step()
                self.pending_begin = True
                return False
            self.index += 1
        return True

    def origin(self):
        """(t0, attempts0) of the current field step: the state's own origin once the step's first
        launch has set it, else the state as it is (the origin that launch will set)."""
        st = self.st
        return (st.t, st.attempts) if self.pending_begin else (st.t0, st.attempts0)

    def step_done(self, st) -> bool:
        """'update': `count` update() calls have run (each steps unconditionally, also once t has
        overflowed to +inf). 'progress': the call has returned, i.e. its predicate (core.py:756)
        fails on the MC steps or on the elapsed time (also an infinite or NaN one). A step whose
        first launch has not run yet is not done."""
        if self.pending_begin:
            return False
        cur = self.current
        if cur.advance == 'update':
            return st.attempts - st.attempts0 >= cur.count
        return not (lower_than(st.t - st.t0, cur.t_max)
                    and lower_than(st.attempts / self.mm.n - st.attempts0 / self.mm.n, cur.MCsteps_max))

    def neel_item(self, chunk, nblocks):
        """(launch item, first_free) of this Neel job's next launch."""
        cur, st, start = self.current, self.st, self.origin()
        if cur.advance == 'update':
            steps = min(cur.count - (st.attempts - start[1]), chunk)
            if self.mm.recompute_every:
                steps = min(steps, max(1, self.mm.recompute_every - self.mm._since_recompute))
            return (self.mm, max(1, steps), cur.t_max, math.inf, nblocks), 2
        steps = self.batch.neel_chunk(self.mm, start, cur.t_max, cur.MCsteps_max, chunk, st)
        return (self.mm, steps, cur.t_max, cur.MCsteps_max, nblocks), 0

    def metro_item(self, chunk):
        cur = self.current
        steps = self.batch.metropolis_chunk(self.mm, self.sel, self.origin(), cur.MCsteps_max, chunk, self.st)
        return (self.mm, self.sel, steps, cur.t_max, cur.MCsteps_max)


class Server:
    """Continuous-batching loop. `budget_bytes` bounds the summed footprint of admitted jobs;
    `max_workgroups` bounds one Neel launch (default: one 1024-thread workgroup per CU);
    `metro_chunk` caps the Metropolis steps per iteration so that new arrivals are admitted
    promptly; `overlap` (default) runs the Neel and Metropolis launches of an iteration on two
    streams of the default priority, so their kernels share the device (measured on MI355X: 31% of
    the Neel launches of a 16-job mix overlap Metropolis kernels, 2% more attempts/s, p50 latency
    3% lower).
    `backend` replaces the module of batched launches (hotspice_amd.batch), e.g. by a host
    stand-in that exercises the control flow without a device."""

    def __init__(self, budget_bytes=64 << 30, max_workgroups=None, neel_chunk=1024, metro_chunk=256,
                 clock=time.perf_counter, overlap=True, backend=None):
        import torch
        self.torch = torch
        self.batch = batch if backend is None else backend
        self.budget = int(budget_bytes)
        if max_workgroups is None:
            max_workgroups = torch.cuda.get_device_properties(0).multi_processor_count
        self.max_wg = int(max_workgroups)
        self.neel_chunk = int(neel_chunk)
        self.metro_chunk = int(metro_chunk)
        self.clock = clock
        self.overlap = bool(overlap) and backend is None  # a stand-in backend launches nothing to overlap
        self._streams = (torch.cuda.Stream(), torch.cuda.Stream()) if self.overlap else None
        self.queue = []  # heap of (arrival, seq, job)
        self.active: List[_Active] = []
        self.done: List[Job] = []
        self.failed: List[Job] = []
        self.in_use = 0
        self._seq = 0
        self._rr = 0
        self.iterations = 0
        self.attempts = 0

    def submit(self, job: Job):
        heapq.heappush(self.queue, (job.arrival, self._seq, job))
        self._seq += 1

    # ---- job life cycle ---------------------------------------------------------------------
    def _fail(self, job, err, now, mm=None, counted=True):
        job.status = "failed"
        job.error = f"{type(err).__name__}: {err}" if isinstance(err, BaseException) else str(err)
        job.finished = now
        if counted:
            self.in_use -= job.footprint
        if mm is not None:
            try:
                mm.release_device()
            except Exception:  # noqa: BLE001 -- the job has failed already; keep serving the others
                pass
        self.failed.append(job)

    def _admit(self, now):
        while self.queue and self.queue[0][0] <= now:
            job = self.queue[0][2]
            if self.active and self.in_use + job.footprint > self.budget:
                break  # wait until memory frees up (a lone job is always admitted)
            heapq.heappop(self.queue)
            mm = None
            try:
                mm = job.make()
                a = _Active(job, mm, now, self.batch)
                complete = a.begin()
                a.st = self.batch.read_states([mm])[0]
            except Exception as e:  # noqa: BLE001 -- one job's error must not stop the others
                self._fail(job, e, now, mm, counted=False)
                continue
            self.in_use += job.footprint
            if complete:
                self._finish(a, now)
            else:
                self.active.append(a)

    def _finish(self, a, now):
        a.mm._after_device_step()
        a.job.finished = now
        a.job.status = "done"
        a.job.result = dict(m=a.mm.m.copy(), t=a.mm.t, attempts=a.mm.attempted_switches, switches=a.mm.switches,
                            E_tot=a.mm.E_tot, n=a.mm.n)
        self.in_use -= a.job.footprint
        self.done.append(a.job)
        a.mm.release_device()

    # ---- one iteration ------------------------------------------------------------------------
    def _launch(self, group, enqueue):
        """Enqueues one group's launch; a descriptor the library rejects fails every job of the
        launch that is rejected on its own (the others are stepped by a launch without them)."""
        items, acts = group
        try:
            enqueue(items)
            return acts
        except _lib.HotspiceError as e:
            if len(items) == 1:
                self._pending_fail.append((acts[0], e))
                return []
        # find the jobs the library rejects: each is enqueued alone (the library checks every
        # descriptor before it launches anything, so a rejected call stepped no job; an accepted
        # job steps alone exactly as it would in the batch)
        good = []
        for it, a in zip(items, acts):
            try:
                enqueue([it])
                good.append(a)
            except _lib.HotspiceError as e:
                self._pending_fail.append((a, e))
        return good

    def _iteration(self):
        """One launch per scheme over the active jobs, sized from the states the last collect read
        back (no read before the launches), then one collect of the launched jobs' states."""
        neel = [a for a in self.active if a.scheme == Scheme.NEEL]
        metro = [a for a in self.active if a.scheme == Scheme.METROPOLIS]
        self._pending_fail = []
        bt = self.batch
        groups = []
        if neel:
            # round-robin packing of teams up to the co-resident workgroup limit
            order = neel[self._rr % len(neel):] + neel[:self._rr % len(neel)]
            self._rr += 1
            items, frees, acts, used = [], [], [], 0
            for a in order:
                g = bt.neel_team_size(a.mm)
                if items and used + g > self.max_wg:
                    continue
                it, free = a.neel_item(self.neel_chunk, g)
                items.append(it)
                frees.append(free)
                acts.append(a)
                used += g
            free_of = {id(it[0]): f for it, f in zip(items, frees)}
            begin_of = {id(a.mm): int(a.pending_begin) for a in acts}
            groups.append(("neel", (items, acts),
                           lambda its: bt.enqueue_neel(its, 1e10, free=[free_of[id(i[0])] for i in its],
                                                       begin=[begin_of[id(i[0])] for i in its])))
        if metro:
            items = [a.metro_item(self.metro_chunk) for a in metro]
            f0 = metro[0].job.kwargs.get('attempt_freq', 1e10)
            mbegin_of = {id(a.mm): int(a.pending_begin) for a in metro}
            groups.append(("metropolis", (items, metro),
                           lambda its: bt.enqueue_metropolis(its, f0, begin=[mbegin_of[id(i[0])] for i in its])))
        stepped = []
        if self.overlap and len(groups) == 2:
            torch = self.torch
            main = torch.cuda.current_stream()
            for (what, group, enqueue), s in zip(groups, self._streams):
                s.wait_stream(main)  # the field swaps queued before the launch
                with torch.cuda.stream(s):
                    stepped.append(self._launch(group, enqueue))
            for s in self._streams:
                main.wait_stream(s)
        else:
            for what, group, enqueue in groups:
                stepped.append(self._launch(group, enqueue))
        acts = [a for g in stepped for a in g]
        for a in acts:
            a.pending_begin = False  # its launch has set the field step's origin
        before = [a.st.attempts for a in acts]
        after, errs = bt.collect([a.mm for a in acts], before)
        now = self.clock() - self._t0
        for i, (a, st) in enumerate(zip(acts, after)):
            a.st = st
            if i in errs:
                self._pending_fail.append((a, f"device step error (err={errs[i]})"))
            else:
                self.attempts += st.attempts - before[i]
        for a, e in self._pending_fail:
            self.active.remove(a)
            self._fail(a.job, e, now, a.mm)
        return dict(zip((id(a) for a in acts), after))

    def _retire(self, now, sts):
        keep = []
        for a in self.active:
            st = sts.get(id(a))
            finished = False
            if st is not None:
                try:
                    while a.step_done(st):
                        a.index += 1
                        if a.begin():
                            finished = True
                            break
                except Exception as e:  # noqa: BLE001
                    self._fail(a.job, e, now, a.mm)
                    continue
            if finished:
                self._finish(a, now)
            else:
                keep.append(a)
        self.active = keep

    def run(self, t0=None, timeout=None):
        """Serves until every submitted job is done or failed. Arrival times are seconds after
        `t0` (default: now). Returns the finished jobs (failed ones are in self.failed)."""
        self._t0 = self.clock() if t0 is None else t0
        while self.queue or self.active:
            now = self.clock() - self._t0
            self._admit(now)
            if not self.active:
                if self.queue:
                    wait = self.queue[0][0] - now
                    if wait > 0:
                        time.sleep(min(wait, 0.05))
                continue
            sts = self._iteration()
            self.iterations += 1
            self._retire(self.clock() - self._t0, sts)
            if timeout is not None and self.clock() - self._t0 > timeout:
                raise TimeoutError(f"serving did not finish within {timeout} s")
        return self.done


def latency_stats(jobs):
    lat = np.array([j.finished - j.arrival for j in jobs])
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()),
                max=float(lat.max()))
