"""Field-step drivers and readouts (the interfaces of hotspice/io.py): the inputters are the
reference's sequences of set_field and progress, and the device readouts (hs_binned_sums) equal
the reference's NumPy sums on the downloaded spins."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pinwheel(n=24, seed=5, scheme='Néel'):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    return hs.ASI.IP_Pinwheel(1e-6, n, T=300, E_B=60 * kT * np.exp(0.05 * rng.normal(size=(n, n))), moment=2.37e-16,
                              pattern=np.where(rng.random((n, n)) < .5, -1., 1.), seed=seed,
                              energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(0., 0.)],
                              params=hs.SimParams(UPDATE_SCHEME=scheme))


def _oop(n=20, seed=6):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * 300
    return hs.ASI.OOP_Square(1e-6, n, T=300, E_B=20 * kT, moment=1.1e-15, pattern=np.where(rng.random((n, n)) < .5, -1., 1.),
                             seed=seed, energies=[hs.DipolarEnergy(), hs.ZeemanEnergy(0., 0.)],
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis'))


def _state(mm):
    return np.asarray(mm.m).copy(), mm.E.copy(), mm.t, mm.attempted_switches, mm.switches


def _same(a, b):
    for x, y in zip(a, b):
        if isinstance(x, np.ndarray):
            np.testing.assert_array_equal(x, y)
        else:
            assert x == y


@pytest.mark.parametrize("make", [_pinwheel, _oop])
def test_m_avg_on_device_equals_numpy(make):
    """core.py:413-421 on the downloaded spins vs the device sums."""
    mm = make()
    mm.progress(t_max=math.inf, MCsteps_max=0.5)
    m = np.asarray(mm.m)
    if mm.in_plane:
        mx = float(np.mean(m * mm.orientation[:, :, 0])) * mm.nx * mm.ny / mm.n
        my = float(np.mean(m * mm.orientation[:, :, 1])) * mm.nx * mm.ny / mm.n
        assert mm.m_avg_x == pytest.approx(mx, rel=1e-12, abs=1e-15)
        assert mm.m_avg_y == pytest.approx(my, rel=1e-12, abs=1e-15)
        assert mm.m_avg == pytest.approx(math.hypot(mx, my), rel=1e-12, abs=1e-15)
    else:
        assert mm.m_avg == pytest.approx(float(np.mean(m)), rel=1e-12, abs=1e-15)
    assert mm.m_avg == mm.m_avg  # the device sum is the same on every call


def test_m_avg_after_host_edit():
    mm = _oop()
    m = np.asarray(mm.m).copy()
    m[:] = 1.0
    mm.m = m
    assert mm.m_avg == pytest.approx(1.0)


@pytest.mark.parametrize("nx,ny", [(1, 1), (3, 2), (4, 4), (5, 7)])
def test_regional_reader_equals_numpy(nx, ny):
    """RegionalOutputReader.read_state (io.py:359-377): device sums vs the reference's loops."""
    from hotspice_amd import io
    mm = _pinwheel(n=28)
    mm.progress(t_max=math.inf, MCsteps_max=0.3)
    rd = io.RegionalOutputReader(nx, ny, mm)
    dev = rd.read_state().copy()
    host = rd.read_state(m=mm.m).copy()
    m = np.asarray(mm.m) * np.asarray(mm.moment)
    mxx, myy = m * mm.orientation[:, :, 0], m * mm.orientation[:, :, 1]
    ref = np.zeros(2 * nx * ny)
    for r in range(nx * ny):
        ref[r] = np.sum(mxx[rd.region == r])
        ref[r + nx * ny] = np.sum(myy[rd.region == r])
    ref /= rd.normalization_factor
    np.testing.assert_allclose(dev, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(host, ref, rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(dev, rd.read_state())


def test_chess_reader_equals_numpy():
    """OOPSquareChessOutputReader.read_state (io.py:438-445)."""
    from hotspice_amd import io
    mm = _oop(n=30)
    mm.progress(t_max=math.inf, MCsteps_max=1.0)
    rd = io.OOPSquareChessOutputReader(3, 5, mm)
    dev = rd.read_state().copy()
    masked = rd.AFM_mask * np.asarray(mm.m)
    ref = np.asarray([np.sum(masked[rd.region == r]) for r in range(rd.n)]) / rd.normalization_factor
    np.testing.assert_allclose(dev, ref, rtol=1e-12, atol=1e-15)
    assert rd.n == 15 and rd.node_coords.shape == (15, 2)


def test_clocking_inputter_is_the_reference_sequence():
    """ClockingFieldInputter (examples/Clocking_IP_Pinwheel.py:43-55) equals its set_field /
    progress sequence applied by hand to a twin instance."""
    from hotspice_amd import io
    a, b = _pinwheel(), _pinwheel()
    inp = io.ClockingFieldInputter(io.RandomBinaryDatastream(), magnitude=1.5e-3, angle=math.pi / 4, n=0.5,
                                   frequency=1e3)
    bits = [0, 1, 1, 0]
    inp.input(a, bits)
    Z = b.get_energy('Zeeman')
    for bit in bits:
        for ang in inp.bit_to_angles(bit):
            Z.set_field(magnitude=1.5e-3, angle=ang)
            b.progress(t_max=0.5 / 1e3, MCsteps_max=0.5)
    _same(_state(a), _state(b))
    assert a.switches > 0


def test_field_inputter_full_period():
    from hotspice_amd import io
    a, b = _pinwheel(seed=8), _pinwheel(seed=8)
    inp = io.FieldInputter(io.ConstantDatastream(1.0), magnitude=2e-3, angle=0.3, n=0.4, frequency=1e3,
                           half_period=False)
    inp.input(a, [1.0, 0.5])
    Z = b.get_energy('Zeeman')
    Z.set_field(angle=0.3)
    for v in [1.0, 0.5]:
        Z.set_field(magnitude=2e-3 * v)
        b.progress(t_max=1 / 1e3 / 2, MCsteps_max=0.4 / 2)
        Z.set_field(magnitude=-2e-3 * v)
        b.progress(t_max=.5 / 1e3, MCsteps_max=0.4 / 2)
    _same(_state(a), _state(b))


def test_chess_steps_inputter_metropolis():
    """OOPSquareChessStepsInputter (io.py:468-481): two half-checkerboard field steps per input."""
    from hotspice_amd import io
    a, b = _oop(seed=9), _oop(seed=9)
    inp = io.OOPSquareChessStepsInputter(io.RandomBinaryDatastream(), magnitude=1e-3, n=1.0)
    inp.input(a, [1, 0])
    Z = b.get_energy('Zeeman')
    mask = ((b.ixx + b.iyy) % 2).astype(float)
    for bit in [1, 0]:
        mag = (2 * bit - 1) * 1e-3
        Z.set_field(magnitude=mask * mag)
        b.progress(t_max=1., MCsteps_max=1.0)
        Z.set_field(magnitude=(mask - 1) * mag)
        b.progress(t_max=1., MCsteps_max=1.0)
    _same(_state(a), _state(b))


def test_perp_inputter_relax_minimizes_under_each_field():
    """PerpFieldInputter(relax=True) (io.py:280-284): a positive then a negative field at the bit's
    angle, each followed by Magnets.minimize, equal to that sequence applied by hand to a twin
    with the same numpy.random draws."""
    from hotspice_amd import io
    a, b = _pinwheel(n=20, seed=10), _pinwheel(n=20, seed=10)
    inp = io.PerpFieldInputter(io.RandomBinaryDatastream(), magnitude=3e-3, angle=0.2, relax=True, frequency=0)
    np.random.seed(1)
    inp.input(a, [0, 1])
    np.random.seed(1)
    Z = b.get_energy('Zeeman')
    for bit in [0, 1]:
        for mag in (3e-3, -3e-3):
            Z.set_field(magnitude=mag, angle=0.2 + bit * math.pi / 2)
            b.minimize()
    _same(_state(a), _state(b))
    assert a.switches > 0


def _oracle_twin(mm, scheme):
    """The FP64 oracle instance of a test lattice built above (same pattern, E_B, seed)."""
    from oracle import model
    lat = type(mm).__name__
    return model.make_magnets(lat, 1e-6, n=mm.nx, T=300, E_B=np.asarray(mm.E_B), moment=float(np.max(mm.moment)),
                              pattern=np.asarray(mm.m), seed=mm.seed,
                              energies=[model.Dipolar(), model.Zeeman(0.0, 0.0)],
                              params=dict(UPDATE_SCHEME="Neel" if scheme == "Néel" else "Metropolis",
                                          REDUCED_KERNEL_SIZE=mm.params.REDUCED_KERNEL_SIZE))


def _play_on_oracle(om, steps):
    """io.run_program's semantics on the oracle: set_field, then the step's advance."""
    z = om.get_energy('zeeman')
    for st in steps:
        z.set_field(st.magnitude, st.angle)
        if st.advance == 'progress':
            om.progress(t_max=st.t_max, MCsteps_max=st.MCsteps_max)
        elif st.advance == 'update':
            for _ in range(st.count):
                om.update(t_max=st.t_max)


@pytest.mark.parametrize("which", ["clocking", "field_full", "chess_steps"])
def test_inputter_programs_match_oracle(which):
    """Field programs of the inputters (io.py:221-239, 468-481; examples/Clocking_IP_Pinwheel.py:
    43-55) played on the device (io.run_program) and on the FP64 oracle: after every input the
    same spins and counters, fields within 1e-5 of their scale, time within 1e-4 relative."""
    from hotspice_amd import io
    if which == "chess_steps":
        mm = _oop(n=24, seed=12)
        inp = io.OOPSquareChessStepsInputter(io.RandomBinaryDatastream(), magnitude=2e-3, n=0.5)
        scheme = "Metropolis"
    else:
        mm = _pinwheel(n=20, seed=11)
        scheme = "Néel"
        inp = io.ClockingFieldInputter(io.RandomBinaryDatastream(), magnitude=1.5e-3, angle=math.pi / 4, n=0.5,
                                       frequency=1e6) if which == "clocking" else \
            io.FieldInputter(io.ConstantDatastream(1.0), magnitude=2e-3, angle=0.3, n=0.3, frequency=1e6,
                             half_period=False)
    om = _oracle_twin(mm, scheme)
    scale = float(np.abs(om.E).max())
    for v in [0, 1, 1, 0, 1]:
        steps = inp.program(mm, v)
        io.run_program(mm, steps)
        _play_on_oracle(om, steps)
        np.testing.assert_array_equal(mm.m, om.m)
        assert mm.attempted_switches == om.attempted_switches and mm.switches == om.switches
        np.testing.assert_allclose(mm.E, om.E, rtol=0, atol=1e-5 * scale)
        assert mm.t == pytest.approx(om.t, rel=1e-4)
    assert mm.switches > 0
