"""Field-step inputs and regional readouts around the stepping path, with the interface of
hotspice/io.py (datastreams io.py:20-190, inputters io.py:75-103, 192-308, 384-500, readers
io.py:105-135, 311-445) and the clocking inputter of examples/Clocking_IP_Pinwheel.py:28-55.

Inputs are field programs. An inputter does not step a simulation itself: `program(mm, value)`
compiles one input value into a sequence of FieldStep records, each one ZeemanEnergy.set_field
call (a uniform magnitude at an angle, or a per-cell magnitude array) followed by how the
instance advances under that field (a progress() call, `count` update() calls, a minimize(), or
nothing). `run_program` plays a program through the public API only, where set_field swaps the
Zeeman term of the combined device field in place and progress / update / minimize run on the
device. Because a program is plain data, the same steps can be handed to the continuous-batching
server (serve.Job(program=...)), which advances many instances' programs together. The one
program that depends on how far the instance got -- FieldInputter's sine mode, which re-sets the
magnitude before every update -- is a generator that computes each step when it is reached.

Readouts are binned sums. Every regional reader sums m times a per-cell weight over a grid of
rectangular bins and divides by a per-bin norm; `_BinnedReadout` keeps the bins and weights on
the device (hs_binned_sums), so one read downloads one number per bin instead of the spins.
"""
import dataclasses
import math
import warnings
from abc import ABC, abstractmethod
from inspect import isgeneratorfunction
from typing import Iterable, Iterator, Optional

import numpy as np

from .ASI import OOP_Square
from .core import Magnets, Scheme
from .energies import ZeemanEnergy
from .utils import lower_than


# ---- datastreams (io.py:20-190) --------------------------------------------------------------------
class Datastream(ABC):
    """A source of input values. `rng` starts from OS entropy; reset_rng() restarts the stream
    from seed 0, so it yields the same values again (io.py:20-27)."""

    _param_names = ()

    def __init__(self):
        self.rng = np.random.default_rng(None)

    def reset_rng(self):
        self.rng = np.random.default_rng(0)

    @abstractmethod
    def get_next(self, n=1) -> np.ndarray:
        """Exactly `n` values."""

    @property
    @abstractmethod
    def dtype(self):
        return None

    def __repr__(self):
        args = ", ".join(f"{k}={getattr(self, k)!r}" for k in self._param_names)
        return f"hotspice.io.{type(self).__name__}({args})"


class ScalarDatastream(Datastream):
    dtype = float


class BinaryDatastream(Datastream):
    """Values are 0 or 1."""
    dtype = int


class IntegerDatastream(Datastream):
    """Values are integers of `bits_per_int` bits; inputters apply them bit by bit."""
    dtype = int

    @property
    @abstractmethod
    def bits_per_int(self) -> int:
        pass

    def as_bits(self, integer: int, endianness='little') -> np.ndarray:
        """The `bits_per_int` bits of `integer`, least significant first ('little', the default:
        nearby integers then share their last bits) or most significant first ('big')."""
        if endianness not in ('little', 'big'):
            raise ValueError(f"endianness must be 'little' or 'big', not {endianness!r}")
        width = max(self.bits_per_int, int(integer).bit_length())
        bits = (int(integer) >> np.arange(width)) & 1
        return bits if endianness == 'little' else bits[::-1]


class ConstantDatastream(ScalarDatastream):
    """The value `constant`, always."""
    _param_names = ("constant",)

    def __init__(self, constant: float = 1):
        super().__init__()
        self.constant = constant

    def get_next(self, n=1) -> np.ndarray:
        return np.full(n, self.constant, dtype=float)


class RandomBinaryDatastream(BinaryDatastream):
    """Random bits: 0 with probability `p0`, else 1 (a uniform draw at or above p0 is a 1)."""
    _param_names = ("p0",)

    def __init__(self, p0=.5):
        super().__init__()
        self.p0 = p0

    def get_next(self, n=1) -> np.ndarray:
        return self.rng.random(size=(n,)) >= self.p0


class RandomIntegerDatastream(IntegerDatastream):
    """Random integers, uniform over 0 .. 2**num_bits - 1."""
    _param_names = ("num_bits",)

    def __init__(self, num_bits: int = 8):
        if not isinstance(num_bits, int):
            raise ValueError(f"num_bits must be an int, got {type(num_bits).__name__}")
        super().__init__()
        self.num_bits = num_bits

    @property
    def bits_per_int(self) -> int:
        return self.num_bits

    def get_next(self, n=1) -> np.ndarray:
        return self.rng.integers(0, 1 << self.num_bits, size=(n,))


class RandomScalarDatastream(ScalarDatastream):
    """Random floats, uniform in [low, high)."""
    _param_names = ("low", "high")

    def __init__(self, low=0, high=1):
        super().__init__()
        self.low, self.high = low, high

    def get_next(self, n=1) -> np.ndarray:
        return self.low + (self.high - self.low) * self.rng.random(size=(n,))


# ---- field programs --------------------------------------------------------------------------------
@dataclasses.dataclass(frozen=True)
class FieldStep:
    """One ZeemanEnergy.set_field(magnitude, angle) call (None keeps that part of the field) and
    what follows it under that field:
      'progress'  mm.progress(t_max, MCsteps_max)
      'update'    `count` calls of mm.update(t_max=t_max)
      'minimize'  mm.minimize()
      'none'      nothing (the next step changes the field again)."""
    magnitude: object = None
    angle: Optional[float] = None
    advance: str = 'progress'
    t_max: float = math.inf
    MCsteps_max: float = 4.0
    count: int = 1

    def __post_init__(self):
        if self.advance not in ('progress', 'update', 'minimize', 'none'):
            raise ValueError(f"unknown advance {self.advance!r}")


def zeeman_of(mm: Magnets) -> ZeemanEnergy:
    """The instance's Zeeman energy; one at zero field is added when it has none (io.py:224-225)."""
    z = mm.get_energy('Zeeman', verbose=False)
    if z is None:
        z = ZeemanEnergy(0, 0)
        mm.add_energy(z)
    return z


def run_program(mm: Magnets, steps: Iterable[FieldStep]):
    """Plays field steps on `mm` through the public API, in order."""
    z = zeeman_of(mm)
    for st in steps:
        z.set_field(magnitude=st.magnitude, angle=st.angle)
        if st.advance == 'progress':
            mm.progress(t_max=st.t_max, MCsteps_max=st.MCsteps_max)
        elif st.advance == 'update':
            for _ in range(st.count):
                mm.update(t_max=st.t_max)
        elif st.advance == 'minimize':
            mm.minimize()


def _needs_neel(mm: Magnets, what: str):
    if mm.params.UPDATE_SCHEME != Scheme.NEEL:
        raise AttributeError(f"{what} needs UPDATE_SCHEME = Scheme.NEEL (it advances by time).")


def _checkerboard(mm: Magnets) -> np.ndarray:
    """(x + y) mod 2 over the grid: the two sublattices of the square lattice's AFM order."""
    return ((mm.ixx + mm.iyy) % 2).astype(float)


# ---- inputters (io.py:75-103, 192-308, 384-500) ----------------------------------------------------
class Inputter(ABC):
    """Applies input values to a simulation. Subclasses give the field program of one value
    (`program`); `input_single` runs it. An input_single written as a generator (the reference
    allows that) is run to its end by `input`."""

    def __init__(self, datastream: Datastream = None):
        self.datastream = ConstantDatastream(constant=1) if datastream is None else datastream

    def values_to_inputs(self, values) -> np.ndarray:
        """The scalars that input_single receives for `values`: the values themselves, or the bits
        of each integer (as_bits, little-endian) for an IntegerDatastream."""
        values = np.asarray(values).reshape(-1)
        if isinstance(self.datastream, IntegerDatastream):
            return np.concatenate([self.datastream.as_bits(int(v)) for v in values]) if values.size else values
        return values

    def input(self, mm: Magnets, values=None, remove_stimulus=False) -> np.ndarray:
        """Applies `values` (by default the datastream's next value) one input at a time; returns
        the values. With remove_stimulus the field magnitude is zeroed afterwards."""
        values = np.asarray(self.datastream.get_next(n=1) if values is None else values).reshape(-1)
        generator = isgeneratorfunction(self.input_single)
        for v in self.values_to_inputs(values):
            out = self.input_single(mm, v)
            if generator:
                for _ in out:
                    pass
        if remove_stimulus:
            self.remove_stimulus(mm)
        return values

    def input_single(self, mm: Magnets, value):
        run_program(mm, self.program(mm, value))

    def program(self, mm: Magnets, value) -> Iterable[FieldStep]:
        raise NotImplementedError(f"{type(self).__name__} defines input_single directly and has no field program.")

    def schedule(self, mm: Magnets, values) -> list:
        """The field programs of all `values`, concatenated (for a caller that advances the steps
        itself, e.g. the serving loop). Only for programs that do not depend on the run."""
        steps = []
        for v in self.values_to_inputs(values):
            p = self.program(mm, v)
            if isinstance(p, Iterator):
                raise ValueError(f"{type(self).__name__}'s program depends on the run and cannot be scheduled ahead.")
            steps.extend(p)
        return steps

    def remove_stimulus(self, mm: Magnets):
        mm.get_energy('Zeeman').magnitude = 0


class FieldInputter(Inputter):
    """A field along `angle` of magnitude `magnitude * value`, for at most `n` MC sweeps and
    1/frequency seconds per input (io.py:193-239). half_period=False follows it by the reversed
    field for a second half. sine=True (Neel only) instead re-sets the magnitude along
    sin(2 pi period * progress) before each single update, where progress is the larger of the
    sweeps and the time done, as fractions of n and 1/frequency."""

    def __init__(self, datastream: Datastream, magnitude=1, angle=0, n=2, sine=False, frequency=1, half_period=True):
        super().__init__(datastream)
        self.angle = angle
        self.magnitude = magnitude
        self.n = n
        self.sine = sine
        self.frequency = frequency
        self.half_period = half_period

    @property
    def angle(self):
        return self._angle

    @angle.setter
    def angle(self, value):
        self._angle = value % math.tau

    @property
    def full_period(self):
        return not self.half_period

    def program(self, mm: Magnets, value):
        if self.sine:
            _needs_neel(mm, "FieldInputter(sine=True)")
            return self._sine_program(mm, value)
        # the angle is set on its own first, then the magnitude (two set_field calls, io.py:226-235)
        steps = [FieldStep(angle=self.angle if mm.in_plane else None, advance='none')]
        frac = 0.5 if self.full_period else 1.0  # share of the period under each sign
        steps.append(FieldStep(magnitude=self.magnitude * value, t_max=frac / self.frequency, MCsteps_max=self.n * frac))
        if self.full_period:
            steps.append(FieldStep(magnitude=-self.magnitude * value, t_max=0.5 / self.frequency, MCsteps_max=self.n / 2))
        return steps

    def _sine_program(self, mm: Magnets, value):
        yield FieldStep(angle=self.angle if mm.in_plane else None, advance='none')
        mc0, t0 = mm.MCsteps, mm.t
        cycles = 0.5 if self.half_period else 1.0
        while True:
            done = max((mm.MCsteps - mc0) / self.n, (mm.t - t0) * self.frequency)
            if not lower_than(done, 1):
                return
            # each update covers at most 1/20 of the input's duration
            yield FieldStep(magnitude=self.magnitude * value * math.sin(2 * math.pi * cycles * done), advance='update',
                            t_max=min(1 - done, 0.05) / self.frequency)


class FieldInputterBinary(FieldInputter):
    """FieldInputter over a full period whose magnitude is magnitudes[1] for a 1 and magnitudes[0]
    for a 0 (io.py:242-253)."""

    def __init__(self, datastream: Datastream, magnitudes: tuple = (.8, 1), **kwargs):
        self.magnitude_ratio = magnitudes[0] / magnitudes[1]
        super().__init__(datastream, magnitude=magnitudes[1], half_period=False, **kwargs)

    def program(self, mm: Magnets, value):
        return super().program(mm, 1 if value else self.magnitude_ratio)


class PerpFieldInputter(FieldInputter):
    """In-plane only: a field along `angle` for a 0 and along angle + pi/2 for a 1 (intermediate
    values give intermediate angles), applied with magnitude +magnitude then -magnitude. With
    relax=True each of the two fields is followed by minimize(), otherwise by progress() for
    half a period (io.py:259-288)."""

    def __init__(self, datastream: BinaryDatastream, magnitude=1, angle=0, n=2, relax=True, frequency=1, **kwargs):
        self.relax = relax
        super().__init__(datastream, magnitude=magnitude, angle=angle, n=n, frequency=frequency, **kwargs)

    def _check(self, mm):
        if (self.sine or self.frequency) and mm.params.UPDATE_SCHEME != Scheme.NEEL:
            raise AttributeError("PerpFieldInputter with sine=True or a nonzero frequency needs UPDATE_SCHEME = Scheme.NEEL.")
        if not mm.in_plane:
            raise AttributeError("PerpFieldInputter needs an in-plane ASI.")

    def program(self, mm: Magnets, value):
        self._check(mm)
        angle = self.angle + value * math.pi / 2
        if self.relax:
            return [FieldStep(magnitude=s * self.magnitude, angle=angle, advance='minimize') for s in (1, -1)]
        return [FieldStep(magnitude=s * self.magnitude, angle=angle, t_max=0.5 / self.frequency, MCsteps_max=self.n)
                for s in (1, -1)]

    def input_single_generalized(self, mm: Magnets, value):
        """One field along the bit's angle, for 1/frequency seconds and n sweeps; with sine=True
        its magnitude follows a sine (a half or a full period) re-set before each update, at
        least 8 updates per period (io.py:290-308)."""
        self._check(mm)
        run_program(mm, self._generalized_program(mm, value))

    def _generalized_program(self, mm, value):
        angle = self.angle + value * math.pi / 2
        if not self.sine:
            yield FieldStep(magnitude=self.magnitude, angle=angle, t_max=1 / self.frequency, MCsteps_max=self.n)
            return
        mc0, t0 = mm.MCsteps, mm.t
        while True:
            done = max((mm.MCsteps - mc0) / self.n, (mm.t - t0) * self.frequency)
            if not lower_than(done, 1):
                return
            yield FieldStep(magnitude=self.magnitude * math.sin(done * math.pi * (2 - self.half_period)), angle=angle,
                            advance='update', t_max=min(1 - done, 0.125) / self.frequency)


class ClockingFieldInputter(FieldInputter):
    """Clocking of in-plane ASI (examples/Clocking_IP_Pinwheel.py:28-55): for a 0, a field at
    angle + spread for half a period, then at angle - spread; for a 1 the same two fields turned
    by 180 degrees. Each half addresses one sublattice."""

    def __init__(self, datastream: BinaryDatastream, magnitude=1, angle=0, spread=math.pi / 8, n=2, frequency=1):
        self.spread = spread
        super().__init__(datastream, magnitude=magnitude, angle=angle, n=n, frequency=frequency)

    def bit_to_angles(self, bit):
        first, second = self.angle + self.spread, self.angle - self.spread
        return (first + math.pi, second + math.pi) if bit else (first, second)

    def program(self, mm: Magnets, value):
        if self.frequency and mm.params.UPDATE_SCHEME != Scheme.NEEL:
            raise AttributeError("ClockingFieldInputter with a nonzero frequency needs UPDATE_SCHEME = Scheme.NEEL.")
        if not mm.in_plane:
            raise AttributeError("ClockingFieldInputter needs an in-plane ASI.")
        return [FieldStep(magnitude=self.magnitude, angle=a, t_max=0.5 / self.frequency, MCsteps_max=self.n)
                for a in self.bit_to_angles(value)]


class OOPSquareChessFieldInputter(Inputter):
    """Out-of-plane square ASI: a checkerboard field of +-magnitude*(2*value - 1), which favours
    one of the two AFM ground states, for 1/frequency seconds and at most n sweeps
    (io.py:385-404)."""

    def __init__(self, datastream: Datastream, magnitude: float = 1, n=4, frequency=1):
        super().__init__(datastream)
        self.magnitude = magnitude
        self.n = n
        self.frequency = frequency

    def program(self, mm: OOP_Square, value):
        sign = 2 * _checkerboard(mm) - 1
        return [FieldStep(magnitude=sign * (2 * value - 1) * self.magnitude, t_max=1 / self.frequency,
                          MCsteps_max=self.n)]


class OOPSquareChessStepsInputter(Inputter):
    """The checkerboard input in two steps, one half of the checkerboard at a time, so that domain
    walls do not run to saturation (io.py:448-481): the field `transform(value)` on the cells with
    (x + y) odd, then its negative on the others, each for 1/frequency seconds and n sweeps.
    transform maps bits to -+magnitude and scalars piecewise linearly through (0, -magnitude),
    (s, -magnitude + magnitude_range), (1 - s, magnitude - magnitude_range) and (1, magnitude),
    with s = (1 - transition_range)/2."""

    def __init__(self, datastream: Datastream, magnitude: float = 1., magnitude_range: float = 0.,
                 transition_range: float = 1., n: float = 4., frequency: float = 1.):
        super().__init__(datastream)
        self.magnitude, self.magnitude_range, self.transition_range = magnitude, magnitude_range, transition_range
        self.n = n
        self.frequency = frequency
        self.transform = self._get_transform_func()

    def program(self, mm: OOP_Square, value):
        mag = self.transform(value)
        odd = _checkerboard(mm)
        return [FieldStep(magnitude=odd * mag, t_max=1 / self.frequency, MCsteps_max=self.n),
                FieldStep(magnitude=(odd - 1) * mag, t_max=1 / self.frequency, MCsteps_max=self.n)]

    def _get_transform_func(self):
        if isinstance(self.datastream, (BinaryDatastream, IntegerDatastream)):
            return lambda bit: self.magnitude * (2 * bit - 1)
        if isinstance(self.datastream, ScalarDatastream):
            return self._val_to_mag_piecewiselinear

    def _val_to_mag_piecewiselinear(self, value: float):
        M, R, W = self.magnitude, self.magnitude_range, self.transition_range
        s = 0.5 - W / 2
        outer_slope = R / s  # the two outer segments rise by R over a width s
        if value < s:
            return -M + value * outer_slope
        if value > 1 - s:
            return M + (value - 1) * outer_slope
        return 2 * (M - R) / W * (value - 0.5)  # the middle segment, through zero at 0.5


# ---- output readers (io.py:105-135, 311-445) -------------------------------------------------------
class OutputReader(ABC):
    """Reads a state vector from a simulation. configure_for(mm) prepares the reader for one
    instance (done by the constructor when mm is given); read_state sets and returns self.state."""

    def __init__(self, mm: Magnets = None):
        self.mm = mm
        if mm is not None:
            self.configure_for(mm)

    def configure_for(self, mm: Magnets):
        self.mm = mm

    @abstractmethod
    def read_state(self, mm: Magnets = None):
        if mm is not None and mm is not self.mm:
            self.configure_for(mm)

    @property
    def n(self) -> int:
        """Length of the state vector."""
        return self.state.size

    @property
    def node_coords(self) -> np.ndarray:
        """(n, 2) representative x, y of each state entry."""
        return self._node_coords


class FullOutputReader(OutputReader):
    """The spins of the occupied cells, in row-major order (io.py:311-331)."""

    def configure_for(self, mm: Magnets):
        self.mm = mm
        self.indices = np.nonzero(mm.occupation)
        self._node_coords = np.stack([mm.xx[self.indices], mm.yy[self.indices]], axis=1)
        self.state = np.asarray(mm.m)[self.indices]

    def read_state(self, mm: Magnets = None, m=None):
        super().read_state(mm)
        self.state = np.asarray(self.mm.m if m is None else m)[self.indices]
        return self.state


def _bins(nbins, ncells):
    """Bin of every cell along one axis -- int(linspace(0, nbins*(1 - 1e-15), ncells)), the
    reference's binning (io.py:347-348), whose bins are contiguous and ascending -- and the first
    cell of each bin (plus ncells at the end)."""
    idx = np.linspace(0, nbins * (1 - 1e-15), ncells).astype(int)
    starts = np.concatenate([np.searchsorted(idx, np.arange(nbins), side='left'), [ncells]])
    return idx, starts


class _BinnedReadout(OutputReader):
    """A state of sums of m * w[k] over nx * ny rectangular regions, divided by a norm per entry:
    entry k*(nx*ny) + r is component k over region r. Subclasses give the weights, the norm and
    what the state holds; the sums run on the device (hs_binned_sums) over bins and weights
    uploaded at configure_for."""

    def _weights(self, mm) -> np.ndarray:
        raise NotImplementedError

    def _norm(self, mm) -> np.ndarray:
        raise NotImplementedError

    def configure_for(self, mm: Magnets):
        self.mm = mm
        ix, self._xs = _bins(self.nx, mm.nx)
        iy, self._ys = _bins(self.ny, mm.ny)
        self.region = iy[:, None] * self.nx + ix[None, :]
        nreg = self.nx * self.ny
        counts = np.bincount(self.region.ravel(), minlength=nreg)
        self._node_coords = np.stack([np.bincount(self.region.ravel(), weights=np.asarray(mm.xx, float).ravel(), minlength=nreg),
                                      np.bincount(self.region.ravel(), weights=np.asarray(mm.yy, float).ravel(), minlength=nreg)],
                                     axis=1) / counts[:, None]
        w = self._weights(mm)
        self._w_host = w
        self._device_bins = mm._dev.bins(self._xs, self._ys, w)
        self.normalization_factor = self._norm(mm)
        self._node_coords = np.tile(self._node_coords, (len(w), 1))
        self.state = np.zeros(len(w) * nreg)

    def read_state(self, mm: Magnets = None, m=None) -> np.ndarray:
        super().read_state(mm)
        if self.mm is None:
            raise AttributeError(f"{type(self).__name__} is not configured for a Magnets object yet.")
        if m is None:
            self.mm._sync_m()
            sums = self.mm._dev.binned_sums(self._device_bins).reshape(-1)
        else:  # an explicit state is summed on the host
            flat = self.region.ravel()
            mv = np.asarray(m, dtype=float).ravel()
            sums = np.concatenate([np.bincount(flat, weights=mv * wk.ravel(), minlength=self.nx * self.ny)
                                   for wk in self._w_host])
        self.state = sums / self.normalization_factor
        return self.state


class RegionalOutputReader(_BinnedReadout):
    """The moment-weighted magnetisation of each of nx x ny regions, normalised by the region's
    total moment: x then y components for in-plane ASI, one value per region otherwise
    (io.py:335-377)."""

    def __init__(self, nx: int, ny: int, mm: Magnets = None):
        self.nx, self.ny = nx, ny
        self.n_regions = nx * ny
        super().__init__(mm)

    def _weights(self, mm):
        mu = np.asarray(mm.moment, dtype=float) * np.ones((mm.ny, mm.nx))
        if mm.in_plane:
            return np.stack([mu * mm.orientation[:, :, 0], mu * mm.orientation[:, :, 1]])
        return mu[None]

    def _norm(self, mm):
        mu = np.asarray(mm.moment, dtype=float) * np.ones((mm.ny, mm.nx))
        per_region = np.bincount(self.region.ravel(), weights=(mm.occupation * mu).ravel(), minlength=self.n_regions)
        return np.tile(per_region, 2 if mm.in_plane else 1)


class OOPSquareChessOutputReader(_BinnedReadout):
    """For out-of-plane square ASI: the spins times the checkerboard sign (+-1) averaged over each
    of nx x ny regions, which tells the two AFM ground states and their domain walls apart
    (io.py:404-445)."""

    def __init__(self, nx: int, ny: int, mm: OOP_Square = None):
        self.nx, self.ny = int(nx), int(ny)
        super().__init__(mm)

    def configure_for(self, mm: OOP_Square):
        if not isinstance(mm, OOP_Square):
            if mm.in_plane:
                raise TypeError(f"{type(self).__name__} reads out-of-plane ASI only, not {type(mm).__name__}.")
            warnings.warn(f"{type(self).__name__} is meant for OOP_Square, got {type(mm).__name__}.", stacklevel=2)
        self.AFM_mask = (2 * _checkerboard(mm) - 1).astype(int)
        super().configure_for(mm)

    def _weights(self, mm):
        return self.AFM_mask[None].astype(float)

    def _norm(self, mm):
        return np.bincount(self.region.ravel(), weights=np.asarray(mm.occupation, float).ravel(),
                           minlength=self.nx * self.ny).astype(int)
