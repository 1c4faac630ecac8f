"""Continuous-batching server: every job's result is the same served alone and served in a mix
of jobs that arrive while others run (bit-identical spins, counters, time and energy), and the
server reports latencies for every job."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _jobs():
    from hotspice_amd.workloads import served_jobs
    # small instances and few field steps so the test stays short; arrivals 2 ms apart on average
    return served_jobs(count=8, rate=500.0, seed=7, n_range=(16, 40), field_steps=4, neel_work=150, metro_sweeps=1.0)


def _serve(jobs, **kw):
    from hotspice_amd.serve import Server
    srv = Server(**kw)
    for j in jobs:
        srv.submit(j)
    return srv.run(timeout=300), srv


def _check_served_equals_alone(make_jobs):
    mixed, srv = _serve(make_jobs())
    assert len(mixed) == len(make_jobs()) and srv.iterations > 0
    by_name = {(j.name, j.arrival): j.result for j in mixed}
    for job in make_jobs():
        job.arrival = 0.0
        (alone,), _ = _serve([job])
        match = [r for (name, arr), r in by_name.items() if name == job.name]
        assert any(np.array_equal(r["m"], alone.result["m"]) and r["t"] == alone.result["t"]
                   and r["attempts"] == alone.result["attempts"] and r["switches"] == alone.result["switches"]
                   and r["E_tot"] == alone.result["E_tot"] for r in match), job.name


def test_served_job_equals_alone():
    _check_served_equals_alone(_jobs)


@pytest.mark.parametrize("seed,cutoff,n_range", [(11, 5, (24, 96)), (13, None, (16, 48))])
def test_served_mix_equals_alone(seed, cutoff, n_range):
    """Other mixes: windowed kernels with grids large enough for the presampled Metropolis steps,
    and full kernels at another seed (lattices, schemes, sizes and arrivals drawn from it)."""
    from hotspice_amd.workloads import served_jobs
    _check_served_equals_alone(lambda: served_jobs(count=6, rate=800.0, seed=seed, n_range=n_range, field_steps=3,
                                                   neel_work=200, metro_sweeps=1.0, cutoff_cells=cutoff))


def test_latency_stats_and_budget():
    from hotspice_amd.serve import latency_stats
    jobs = _jobs()
    # a budget that holds about two jobs at a time forces admission to wait for retirements
    budget = 2 * max(j.footprint for j in jobs)
    done, srv = _serve(jobs, budget_bytes=budget)
    assert len(done) == len(jobs)
    st = latency_stats(done)
    assert 0 < st["p50"] <= st["p99"] <= st["max"]
    assert all(j.admitted >= j.arrival - 1e-9 and j.finished >= j.admitted for j in done)


def test_served_jobs_complete_after_time_overflow():
    """Jobs whose time overflows to +inf (2000 kT barriers): a Neel job's field steps are `work`
    update(t_max=inf) calls, which step unconditionally, so it keeps switching with t = +inf and
    ends with the same spins and counters as those calls made through the drop-in API; a
    Metropolis job with METROPOLIS_TIME ends each field step when progress() would return (the
    elapsed time is infinite). Both complete, and neither loops forever."""
    import math

    import hotspice_amd as hs
    from hotspice_amd.serve import Job, Server
    from hotspice_amd.workloads import kB

    def make(scheme, seed):
        def build():
            rng = np.random.default_rng(seed)
            n = 16
            kT = kB * 300
            params = hs.SimParams(UPDATE_SCHEME=scheme, METROPOLIS_TIME=(scheme == "Metropolis"))
            return hs.ASI.IP_Pinwheel(1e-6, n, T=300, E_B=2000 * kT * rng.normal(1, 0.01, size=(n, n)),
                                      moment=2.37e-16, pattern=np.where(rng.random((n, n)) < .5, -1., 1.), seed=seed,
                                      energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(0.0, 0.0)], params=params)
        return build

    fields = [(1e-3, 0.0), (2e-3, math.pi / 4), (1e-3, 0.0)]
    jobs = [Job(make=make("Néel", 1), fields=fields, work=40, name="neel"),
            Job(make=make("Metropolis", 2), fields=fields, work=1.0, name="metro")]
    srv = Server()
    for j in jobs:
        srv.submit(j)
    done = {j.name: j for j in srv.run(timeout=120)}
    assert set(done) == {"neel", "metro"}
    # the Neel job equals its update() calls through the drop-in API
    mm = make("Néel", 1)()
    for mag, ang in fields:
        mm.get_energy('Zeeman').set_field(mag, ang)
        for _ in range(40):
            mm.update(t_max=math.inf)
    r = done["neel"].result
    assert math.isinf(r["t"]) and math.isinf(mm.t)
    assert r["attempts"] == mm.attempted_switches == 120 and r["switches"] == mm.switches
    np.testing.assert_array_equal(r["m"], mm.m)
    # the Metropolis job's first step overflows t; every later field step starts at t = +inf, where
    # progress() takes no step (t - t0 is NaN)
    m = done["metro"].result
    assert math.isinf(m["t"]) and 0 < m["attempts"]
