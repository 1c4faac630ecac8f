"""Statistical parity: observables of the device path against closed forms and against the
FP64 oracle run with NumPy generators (the reference's own random-number structure, no replay),
each over 32 seeds, within 2 sigma of the difference of the means.

  * Metropolis with Zeeman only (independent spins): <m> = tanh(mu B / kT) (Noninteracting.py:55);
  * Neel with equal barriers and no field: waiting times ~ Exp(n f0 e^{-E_B/kT}), winners uniform;
  * square-ice vertex fractions: (1/8, 1/4, 1/2, 1/8) for uncorrelated spins
    (IP_SquareVertices.py:162), type II everywhere for the uniform state;
  * dipolar OOP_Square after a few Metropolis sweeps: magnetisation and NN correlation equal the
    oracle's (NumPy RNG) within 2 sigma.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

kB = 1.380649e-23
SEEDS = range(32)


def test_noninteracting_magnetisation_is_tanh():
    import hotspice_amd as hs
    from hotspice_amd import batch
    T, mu = 300.0, 1e-16
    B = 0.5 * kB * T / mu  # mu B / kT = 0.5
    mags = [hs.ASI.OOP_Square(1e-6, 24, T=T, E_B=0.0, moment=mu, pattern='random', seed=s,
                              energies=[hs.ZeemanEnergy(B)], params=hs.SimParams(UPDATE_SCHEME='Metropolis'))
            for s in SEEDS]
    batch.progress_many(mags, t_max=math.inf, MCsteps_max=10.0)  # equilibrate
    avg = np.zeros(len(mags))
    for _ in range(20):
        batch.progress_many(mags, t_max=math.inf, MCsteps_max=1.0)
        avg += [m.m_avg for m in mags]
    avg /= 20
    expect = math.tanh(0.5)
    sem = avg.std(ddof=1) / math.sqrt(len(avg))
    assert abs(avg.mean() - expect) <= 2 * sem + 1e-3, (avg.mean(), expect, sem)


def test_neel_waiting_times_and_winners():
    import hotspice_amd as hs
    T, EB, f0 = 300.0, 20 * kB * 300.0, 1e10
    mm = hs.ASI.OOP_Square(1e-6, 16, T=T, E_B=EB, moment=1e-16, pattern='random', seed=3,
                           energies=[hs.ZeemanEnergy(0.0)], params=hs.SimParams(UPDATE_SCHEME='Néel'))
    steps = 20000
    wins = np.zeros(mm.ny * mm.nx)
    dt, _ = mm.progress(t_max=math.inf, MCsteps_max=steps / mm.n)
    assert mm.attempted_switches == steps
    rate = mm.n * f0 * math.exp(-EB / (kB * T))
    # mean of `steps` Exp(rate) variables: 1/rate with a standard error of 1/(rate sqrt(steps))
    assert abs(dt / steps * rate - 1) <= 2 / math.sqrt(steps)
    # winners: every cell equally likely
    for _ in range(2000):
        y, x = mm.update(t_max=math.inf)
        wins[int(y) * mm.nx + int(x)] += 1
    mean = wins.mean()
    assert abs(wins.std() / math.sqrt(mean) - 1) < 0.25  # multinomial spread of a uniform draw


def test_vertex_fractions():
    import hotspice_amd as hs
    mm = hs.ASI.IP_Square(1e-6, 200, T=300, E_B=0.0, moment=1e-16, pattern='random', seed=5, PBC=True,
                          energies=[hs.ZeemanEnergy(0.0)], params=hs.SimParams(UPDATE_SCHEME='Metropolis'))
    f = mm.vertex_counts()
    nv = (mm.nx // 2) * (mm.ny // 2)
    for got, p in zip(f, (1 / 8, 1 / 4, 1 / 2, 1 / 8)):
        assert abs(got - p) <= 4 * math.sqrt(p * (1 - p) / nv)
    mm.m = np.where(np.asarray(mm.occupation) != 0, 1.0, 0.0)
    np.testing.assert_array_equal(mm.vertex_counts(), [0, 1, 0, 0])


def test_dipolar_observables_match_oracle_statistics():
    """32 seeds on the device (Philox) and 32 on the FP64 oracle (NumPy generators): the mean
    magnetisation and NN correlation after 3 Metropolis sweeps agree within 2 sigma."""
    import hotspice_amd as hs
    from hotspice_amd import batch
    from oracle import model
    a, n, T, mu = 1e-6, 12, 300.0, 1.1e-15
    kw = dict(r=3e-6)
    dev = [hs.ASI.OOP_Square(a, n, T=T, E_B=0.0, moment=mu, pattern='random', seed=100 + s,
                             energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(2e-5)],
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis')) for s in SEEDS]
    batch.progress_many(dev, t_max=math.inf, MCsteps_max=3.0, **kw)
    ref = []
    for s in SEEDS:
        rng = np.random.default_rng(200 + s)
        om = model.make_magnets("OOP_Square", a, n=n, T=T, E_B=0.0, moment=mu,
                                pattern=np.where(rng.random((n, n)) < .5, -1., 1.),
                                energies=[model.Dipolar(), model.Zeeman(magnitude=2e-5)],
                                params=dict(UPDATE_SCHEME="Metropolis"), seed=300 + s, rng="numpy")
        om.progress(t_max=math.inf, MCsteps_max=3.0, **kw)
        ref.append(om)
    def corr_nn(m):  # <S_i S_j> over the 3x3-cross nearest neighbours, open boundaries (ASI.py:28-36, 74-75)
        p = np.pad(m, 1)
        nsum = p[:-2, 1:-1] + p[2:, 1:-1] + p[1:-1, :-2] + p[1:-1, 2:]
        cnt = np.pad(np.ones_like(m), 1)
        ncnt = cnt[:-2, 1:-1] + cnt[2:, 1:-1] + cnt[1:-1, :-2] + cnt[1:-1, 2:]
        return float(np.mean(m * nsum / ncnt))

    for name, f in [("m_avg", lambda m: float(np.mean(m))), ("correlation_NN", corr_nn)]:
        g = np.array([f(np.asarray(m.m)) for m in dev])
        o = np.array([f(np.asarray(m.m)) for m in ref])
        sig = math.sqrt(g.var(ddof=1) / len(g) + o.var(ddof=1) / len(o))
        assert abs(g.mean() - o.mean()) <= 2 * sig, (name, g.mean(), o.mean(), sig)
