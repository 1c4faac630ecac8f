"""Continuous-batching server for field-step jobs on one device.

A job is a spin-ice instance driven through a sequence of field steps, the pattern of
FieldInputter.input_single (hotspice/io.py:221-239) and of the clocking examples
(examples/Clocking_IP_Pinwheel.py:28-55): for each step the Zeeman field is set
(ZeemanEnergy.set_field, energies.py:101-114) and the instance is advanced by a fixed amount of
work, either `work` Neel switch attempts (that many update(t_max=inf) calls, core.py:679-709) or
progress(t_max=inf, MCsteps_max=work) under Metropolis (core.py:736-760).

The server
  * admits queued jobs whose arrival time has passed while their HBM footprint fits the budget
    (state arrays + dipolar kernel tables, estimated before the instance is built);
  * keeps every admitted job's state resident on the device between iterations;
  * advances all active jobs together in each iteration: one persistent Neel launch (teams
    packed up to the co-resident workgroup limit, the rest wait for the next iteration) and one
    batched Metropolis call, each job running at most to the end of its current field step;
  * admits jobs that arrived meanwhile at the next iteration (continuous batching), frees a job's
    device memory when its last field step is done and records its latency.

Every launch of a job is chunked from the job's own state only (its field-step work, its recompute
period), and the kernels give an instance the same arithmetic alone or in a batch, so a job's
result does not depend on what else is being served.
"""
import dataclasses
import heapq
import math
import time
from typing import Callable, List, Optional, Tuple

import numpy as np

from . import batch
from .core import Scheme
from .utils import lower_than


@dataclasses.dataclass
class Job:
    """One request. `make()` builds its Magnets (on the device); `fields` is the list of
    (magnitude, angle) field steps; `work` is Neel attempts or Metropolis MC sweeps per step."""
    make: Callable
    fields: List[Tuple[float, float]]
    work: float
    arrival: float = 0.0
    footprint: int = 0
    name: str = ""
    kwargs: dict = dataclasses.field(default_factory=dict)  # Metropolis r / Q, attempt_freq
    spec: Optional[dict] = None  # what make() builds (for the caller's records; unused here)
    # filled in by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    result: Optional[dict] = None


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


class _Active:
    def __init__(self, job, mm, now):
        self.job = job
        self.mm = mm
        self.step = 0
        self.scheme = mm.params.UPDATE_SCHEME
        self.sel = None
        if self.scheme == Scheme.METROPOLIS:
            r = job.kwargs.get('r', 0)
            Q = job.kwargs.get('Q', 0.05)
            self.sel = mm._select_geometry(mm.calc_r(Q) if r == 0 else r)
        self.zeeman = mm.get_energy('Zeeman')
        self._begin_step()
        job.admitted = now

    def _begin_step(self):
        mag, ang = self.job.fields[self.step]
        self.zeeman.set_field(magnitude=mag, angle=ang)
        self.mm._pre_step()
        st = self.mm._state()
        self.start = (st.t, st.attempts)
        self.mm._set_progress_origin(*self.start)

    def step_done(self, st=None):
        """Neel: `work` update(t_max=inf) calls have run (each steps unconditionally, also once t
        has overflowed to +inf). Metropolis: progress(t_max=inf, MCsteps_max=work) has returned,
        i.e. its predicate (core.py:756) fails, on the MC steps or on an infinite (or NaN) elapsed
        time from METROPOLIS_TIME."""
        st = self.mm._state() if st is None else st
        if self.scheme == Scheme.NEEL:
            return st.attempts - self.start[1] >= int(self.job.work)
        return not (lower_than(st.t - self.start[0], math.inf)
                    and lower_than(st.attempts / self.mm.n - self.start[1] / self.mm.n, self.job.work))

    def advance(self):
        """Moves to the next field step; returns True when the job is complete."""
        self.step += 1
        if self.step >= len(self.job.fields):
            return True
        self._begin_step()
        return False


class Server:
    """Continuous-batching loop. `budget_bytes` bounds the summed footprint of admitted jobs;
    `max_workgroups` bounds one Neel launch (default: one 1024-thread workgroup per CU);
    `metro_chunk` caps the Metropolis steps per iteration so that new arrivals are admitted
    promptly."""

    def __init__(self, budget_bytes=64 << 30, max_workgroups=None, neel_chunk=1024, metro_chunk=256,
                 clock=time.perf_counter):
        import torch
        self.budget = int(budget_bytes)
        if max_workgroups is None:
            max_workgroups = torch.cuda.get_device_properties(0).multi_processor_count
        self.max_wg = int(max_workgroups)
        self.neel_chunk = int(neel_chunk)
        self.metro_chunk = int(metro_chunk)
        self.clock = clock
        self.queue = []  # heap of (arrival, seq, job)
        self.active: List[_Active] = []
        self.done: List[Job] = []
        self.in_use = 0
        self._seq = 0
        self._rr = 0
        self.iterations = 0
        self.attempts = 0

    def submit(self, job: Job):
        heapq.heappush(self.queue, (job.arrival, self._seq, job))
        self._seq += 1

    def _admit(self, now):
        while self.queue and self.queue[0][0] <= now:
            job = self.queue[0][2]
            if self.active and self.in_use + job.footprint > self.budget:
                break  # wait until memory frees up (a lone job is always admitted)
            heapq.heappop(self.queue)
            mm = job.make()
            self.in_use += job.footprint
            self.active.append(_Active(job, mm, now))

    def _iteration(self):
        neel = [a for a in self.active if a.scheme == Scheme.NEEL]
        metro = [a for a in self.active if a.scheme == Scheme.METROPOLIS]
        if neel:
            # round-robin packing of teams up to the co-resident workgroup limit
            order = neel[self._rr % len(neel):] + neel[:self._rr % len(neel)]
            self._rr += 1
            items, used = [], 0
            for a, st in zip(order, batch.read_states([a.mm for a in order])):
                g = batch.neel_team_size(a.mm)
                if items and used + g > self.max_wg:
                    continue
                left = int(a.job.work) - (st.attempts - a.start[1])
                steps = min(left, self.neel_chunk)
                if a.mm.recompute_every:
                    steps = min(steps, max(1, a.mm.recompute_every - a.mm._since_recompute))
                items.append((a.mm, steps, math.inf, math.inf, g))
                used += g
                self.attempts -= st.attempts
            self.attempts += sum(st.attempts for st in batch.launch_neel(items, 1e10, updates=True))
        if metro:
            items = []
            for a, st in zip(metro, batch.read_states([a.mm for a in metro])):
                steps = batch.metropolis_chunk(a.mm, a.sel, a.start, a.job.work, self.metro_chunk, st)
                items.append((a.mm, a.sel, steps, math.inf, a.job.work))
                self.attempts -= st.attempts
            after = batch.launch_metropolis(items, metro[0].job.kwargs.get('attempt_freq', 1e10))
            self.attempts += sum(st.attempts for st in after)

    def _retire(self, now):
        keep = []
        for a, st in zip(self.active, batch.read_states([a.mm for a in self.active])):
            finished = False
            while a.step_done(st):
                if a.advance():
                    finished = True
                    break
                st = a.mm._state()
            if finished:
                a.mm._after_device_step()
                a.job.finished = now
                a.job.result = dict(m=a.mm.m.copy(), t=a.mm.t, attempts=a.mm.attempted_switches,
                                    switches=a.mm.switches, E_tot=a.mm.E_tot)
                self.in_use -= a.job.footprint
                self.done.append(a.job)
                a.mm.release_device()
            else:
                keep.append(a)
        self.active = keep

    def run(self, t0=None, timeout=None):
        """Serves until every submitted job is done. Arrival times are seconds after `t0`
        (default: now). Returns the finished jobs."""
        t0 = self.clock() if t0 is None else t0
        while self.queue or self.active:
            now = self.clock() - t0
            self._admit(now)
            if not self.active:
                wait = self.queue[0][0] - now
                if wait > 0:
                    time.sleep(min(wait, 0.05))
                continue
            self._iteration()
            self.iterations += 1
            self._retire(self.clock() - t0)
            if timeout is not None and self.clock() - t0 > timeout:
                raise TimeoutError(f"serving did not finish within {timeout} s")
        return self.done


def latency_stats(jobs):
    lat = np.array([j.finished - j.arrival for j in jobs])
    return dict(p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)), mean=float(lat.mean()),
                max=float(lat.max()))
