"""Failure isolation of the serving loop, on the host: a stand-in for the batched launches
(Server(backend=...)) drives the server's control flow without a device. One job cannot be
built, one is rejected by the launch's descriptor check (as hs_check_instance rejects a bad
descriptor before anything is launched), one reports a device error word after its launch, one
asks for steps that cannot be served; each is recorded as failed with its error, and every other
job completes with exactly the work it would do served alone (the reference's runner records a
failed job and carries on, ParallelJobs.py:93-95, 108-112)."""
import math
import types

import numpy as np
import pytest

from hotspice_amd import _lib
from hotspice_amd.core import Scheme
from hotspice_amd.io import FieldStep
from hotspice_amd.serve import Job, Server


class FakeZeeman:
    def __init__(self):
        self.calls = []

    def set_field(self, magnitude=None, angle=None):
        self.calls.append((magnitude, angle))


class FakeMagnets:
    """Enough of Magnets for the server: counters that launches advance, a Zeeman energy."""

    def __init__(self, scheme, n=16, per_step=1, reject=False, device_err=0):
        self.params = types.SimpleNamespace(UPDATE_SCHEME=scheme)
        self._uniform_beta = True
        self.n = n
        self.per_step = per_step
        self.reject = reject
        self.device_err = device_err
        self.recompute_every = None
        self._since_recompute = 0
        self.z = FakeZeeman()
        self.st = types.SimpleNamespace(t=0.0, attempts=0, steps_done=0, err=0, t0=0.0, attempts0=0)
        self.released = False

    def get_energy(self, name, verbose=True):
        return self.z if name.lower() == 'zeeman' else None

    def _select_geometry(self, r):
        return types.SimpleNamespace(nsc=4)

    def calc_r(self, Q):
        return 1.0

    def _pre_step(self):
        pass

    def _state(self):
        return types.SimpleNamespace(**vars(self.st))

    def _set_progress_origin(self, t0, a0):
        self.st.t0, self.st.attempts0 = t0, a0

    def step(self, k):
        self.st.attempts += k * self.per_step
        self.st.steps_done = k
        self.st.t += 1e-9 * k
        if self.device_err:
            self.st.err = self.device_err

    # what the server reads when a job is done
    @property
    def m(self):
        return np.zeros(4)

    @property
    def t(self):
        return self.st.t

    @property
    def attempted_switches(self):
        return self.st.attempts

    switches = 0
    E_tot = 0.0

    def _after_device_step(self):
        pass

    def release_device(self):
        self.released = True


class FakeBatch:
    """The launches of hotspice_amd.batch on the host: a call with a rejected descriptor raises
    before stepping anything, as the library's calls do."""

    def read_states(self, mags):
        return [m._state() for m in mags]

    def neel_team_size(self, m, nblocks=0):
        return 1

    def neel_chunk(self, m, start, t_max, mc, chunk=None, st=None):
        return 3

    def metropolis_chunk(self, m, sel, start, mc, chunk=None, st=None):
        return 2

    def _enqueue(self, items, begin):
        if any(it[0].reject for it in items):
            raise _lib.HotspiceError("bad n_occ 0")
        begin = [begin] * len(items) if np.ndim(begin) == 0 else begin
        for it, b in zip(items, begin):
            if b:  # hs_run_t.begin: the launch records the progress origin first
                it[0]._set_progress_origin(it[0].st.t, it[0].st.attempts)
            it[0].step(it[1] if it[0].params.UPDATE_SCHEME == Scheme.NEEL else it[2])

    def enqueue_neel(self, items, attempt_freq=1e10, begin=False, free=None, full=None):
        self._enqueue(items, begin)

    def enqueue_metropolis(self, items, attempt_freq=1e10, begin=False):
        self._enqueue(items, begin)

    def collect(self, mags, before=None, recomputed=()):
        sts = self.read_states(mags)
        return sts, {i: st.err for i, st in enumerate(sts) if st.err}


def _fields():
    return [(1e-3, 0.0), (2e-3, math.pi / 4), (0.5e-3, 0.0)]


def _jobs():
    def broken():
        raise ValueError("Can not initialize: no spins")
    good = [Job(make=lambda: FakeMagnets(Scheme.NEEL), fields=_fields(), work=10, name="neel ok", arrival=0.0),
            Job(make=lambda: FakeMagnets(Scheme.METROPOLIS, per_step=4), fields=_fields(), work=1.0, name="metro ok",
                arrival=0.0),
            Job(make=lambda: FakeMagnets(Scheme.NEEL), program=[FieldStep(magnitude=1e-3, angle=0.2, advance='none'),
                                                              FieldStep(magnitude=2e-3, t_max=5e-9, MCsteps_max=2.0)],
                name="neel program ok", arrival=0.0)]
    bad = [Job(make=broken, fields=_fields(), work=10, name="cannot build"),
           Job(make=lambda: FakeMagnets(Scheme.NEEL, reject=True), fields=_fields(), work=10, name="rejected neel"),
           Job(make=lambda: FakeMagnets(Scheme.METROPOLIS, reject=True), fields=_fields(), work=1.0,
               name="rejected metro"),
           Job(make=lambda: FakeMagnets(Scheme.METROPOLIS, per_step=4, device_err=3), fields=_fields(), work=1.0,
               name="device error"),
           Job(make=lambda: FakeMagnets(Scheme.METROPOLIS), program=[FieldStep(magnitude=1e-3, advance='update')],
               name="unservable")]
    return good, bad


def _serve(jobs):
    srv = Server(max_workgroups=64, backend=FakeBatch())
    for j in jobs:
        srv.submit(j)
    srv.run(timeout=30)
    return srv


def test_failed_jobs_are_isolated():
    good, bad = _jobs()
    srv = _serve(good + bad)
    assert sorted(j.name for j in srv.done) == sorted(j.name for j in good)
    assert sorted(j.name for j in srv.failed) == sorted(j.name for j in bad)
    assert all(j.status == "done" for j in srv.done) and all(j.status == "failed" and j.error for j in srv.failed)
    errors = {j.name: j.error for j in srv.failed}
    assert "no spins" in errors["cannot build"]
    assert "bad n_occ" in errors["rejected neel"] and "bad n_occ" in errors["rejected metro"]
    assert "err=3" in errors["device error"]
    assert "Neel jobs only" in errors["unservable"]
    assert srv.in_use == 0 and not srv.active
    # every good job did exactly the work it does when served alone
    for j in good:
        alone_good, _ = _jobs()
        (alone,) = [a for a in alone_good if a.name == j.name]
        solo = _serve([alone])
        assert solo.done[0].result["attempts"] == j.result["attempts"]
        assert solo.done[0].result["t"] == j.result["t"]


def test_work_per_step_follows_the_program():
    """Neel field steps of `work` update() calls, Metropolis progress(MCsteps_max=work) steps, and
    a program's progress step ending on its time limit."""
    good, _ = _jobs()
    srv = _serve(good)
    res = {j.name: j.result for j in srv.done}
    assert res["neel ok"]["attempts"] == 3 * 10  # 10 update() calls per field step
    assert res["metro ok"]["attempts"] == 3 * 16  # one sweep of n = 16 per field step, 4 per step
    # t_max = 5e-9 ends the progress step after 5 steps of 1e-9 (launched 3 at a time)
    assert res["neel program ok"]["attempts"] == 6
