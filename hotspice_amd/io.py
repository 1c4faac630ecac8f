"""Inputs and outputs around the stepping path: datastreams, field inputters and output readers
with the interfaces of hotspice/io.py (and the clocking inputter of
examples/Clocking_IP_Pinwheel.py:28-55).

Inputters drive the device path through the public API only: ZeemanEnergy.set_field (an in-place
swap of the Zeeman term in the combined device field) followed by Magnets.progress / update.
The regional readers sum on the device (hs_binned_sums), so reading a state downloads one number
per region instead of the spin array.
"""
import math
import warnings
from abc import ABC, abstractmethod
from inspect import isgeneratorfunction

import numpy as np

from .ASI import OOP_Square
from .core import Magnets, Scheme
from .energies import ZeemanEnergy
from .utils import lower_than


# ---- datastreams (io.py:20-190) --------------------------------------------------------------------
class Datastream(ABC):
    def __init__(self):
        self.rng = np.random.default_rng(None)  # a random seed from OS entropy (io.py:22)

    def reset_rng(self):
        """Seed 0: the stream yields the same values from the start again (io.py:23-27)."""
        self.rng = np.random.default_rng(0)

    @abstractmethod
    def get_next(self, n=1) -> np.ndarray:
        """An array of exactly `n` scalar values."""

    @property
    @abstractmethod
    def dtype(self):
        return None


class ScalarDatastream(Datastream):
    @property
    def dtype(self):
        return float


class BinaryDatastream(Datastream):
    """A Datastream that yields only 0 or 1."""

    @property
    def dtype(self):
        return int


class IntegerDatastream(Datastream):
    @property
    def dtype(self):
        return int

    @property
    @abstractmethod
    def bits_per_int(self) -> int:
        pass

    def as_bits(self, integer: int, endianness='little') -> np.ndarray:
        """io.py:56-62: the bits of `integer`, little-endian by default."""
        bitstring = bin(integer)[2:].zfill(self.bits_per_int)
        if endianness == 'little':
            bitstring = bitstring[::-1]
        return np.asarray([int(bit) for bit in bitstring])


class ConstantDatastream(ScalarDatastream):
    """Always `constant` (io.py:64-73)."""

    def __init__(self, constant: float = 1):
        super().__init__()
        self.constant = constant

    def get_next(self, n=1) -> np.ndarray:
        return np.ones(n, dtype=float) * self.constant

    def __repr__(self):
        return f"hotspice.io.ConstantDatastream(constant={self.constant})"


class RandomBinaryDatastream(BinaryDatastream):
    """Bits that are 0 with probability `p0` (io.py:147-157)."""

    def __init__(self, p0=.5):
        self.p0 = p0
        super().__init__()

    def __repr__(self):
        return f"hotspice.io.RandomBinaryDatastream(p0={self.p0})"

    def get_next(self, n=1) -> np.ndarray:
        return self.rng.random(size=(n,)) >= self.p0


class RandomIntegerDatastream(IntegerDatastream):
    """Integers uniform in [0, 2**num_bits) (io.py:159-178)."""

    def __init__(self, num_bits: int = 8):
        if not isinstance(num_bits, int):
            raise ValueError("Number of bits per 'byte' in RandomUniformByteDatastream must be of type <int>.")
        self.num_bits = num_bits
        self._max = 2 ** self.num_bits
        super().__init__()

    def __repr__(self):
        return f"hotspice.io.RandomIntegerDatastream(num_bits={self.num_bits})"

    @property
    def bits_per_int(self) -> int:
        return self.num_bits

    def get_next(self, n=1) -> np.ndarray:
        return self.rng.integers(0, self._max, size=(n,))


class RandomScalarDatastream(ScalarDatastream):
    """Floats uniform in [low, high) (io.py:180-190)."""

    def __init__(self, low=0, high=1):
        self.low, self.high = low, high
        super().__init__()

    def __repr__(self):
        return f"hotspice.io.RandomScalarDatastream(low={self.low}, high={self.high})"

    def get_next(self, n=1) -> np.ndarray:
        return (self.high - self.low) * self.rng.random(size=(n,)) + self.low


# ---- inputters (io.py:75-103, 192-308, 384-500) ----------------------------------------------------
def _zeeman(mm: Magnets) -> ZeemanEnergy:
    """The instance's Zeeman energy, added (at zero field) when it has none (io.py:224-225)."""
    Zeeman = mm.get_energy('Zeeman', verbose=False)
    if Zeeman is None:
        mm.add_energy(Zeeman := ZeemanEnergy(0, 0))
    return Zeeman


class Inputter(ABC):
    def __init__(self, datastream: Datastream = None):
        self.datastream = ConstantDatastream(constant=1) if datastream is None else datastream

    def input(self, mm: Magnets, values=None, remove_stimulus=False) -> np.ndarray:
        """io.py:80-99: applies `values` (default: the next value of the datastream) one by one;
        integers of an IntegerDatastream are applied bit by bit."""
        if values is None:
            values = self.datastream.get_next(n=1)
        values = np.asarray(values).reshape(-1)
        inputs = values.copy()
        if isinstance(self.datastream, IntegerDatastream):
            inputs = np.asarray([self.datastream.as_bits(int(value)) for value in values]).reshape(-1)
        input_single_is_generator = isgeneratorfunction(self.input_single)
        for value in inputs:
            result = self.input_single(mm, value)
            if input_single_is_generator:
                for _ in result:
                    pass
        if remove_stimulus:
            self.remove_stimulus(mm)
        return values

    def remove_stimulus(self, mm: Magnets):
        mm.get_energy('Zeeman').magnitude = 0

    @abstractmethod
    def input_single(self, mm: Magnets, value):
        """Applies the stimulus of one scalar `value` to `mm`."""


class FieldInputter(Inputter):
    def __init__(self, datastream: Datastream, magnitude=1, angle=0, n=2, sine=False, frequency=1, half_period=True):
        """io.py:193-206: a field at `angle` of magnitude `magnitude*value`, for at most `n` MC
        steps per input (and 1/frequency seconds); with `sine` the magnitude follows a sine over
        the input (Neel only)."""
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

    def input_single(self, mm: Magnets, value):
        """io.py:221-239"""
        if self.sine and mm.params.UPDATE_SCHEME != Scheme.NEEL:
            raise AttributeError("Can not use temporal sine=True if UPDATE_SCHEME != Scheme.NEEL.")
        Zeeman = _zeeman(mm)
        Zeeman.set_field(angle=(self.angle if mm.in_plane else None))
        MCsteps0, t0 = mm.MCsteps, mm.t
        period = 1 / (1 + self.half_period)
        if self.sine:
            while lower_than(progress := max((mm.MCsteps - MCsteps0) / self.n, (mm.t - t0) * self.frequency), 1):
                Zeeman.set_field(magnitude=self.magnitude * value * math.sin(progress * math.pi * 2 * period))
                mm.update(t_max=min(1 - progress, 0.05) / self.frequency)
        else:
            Zeeman.set_field(magnitude=self.magnitude * value)
            mm.progress(t_max=1 / self.frequency / (2 * period), MCsteps_max=self.n / (2 * period))
            if self.full_period:
                Zeeman.set_field(magnitude=-self.magnitude * value)
                mm.progress(t_max=.5 / self.frequency, MCsteps_max=self.n / 2)


class FieldInputterBinary(FieldInputter):
    def __init__(self, datastream: Datastream, magnitudes: tuple = (.8, 1), **kwargs):
        """io.py:242-253: magnitude magnitudes[0] for bit 0, magnitudes[1] for bit 1, full period."""
        self.magnitude_ratio = magnitudes[0] / magnitudes[1]
        super().__init__(datastream, magnitude=magnitudes[1], half_period=False, **kwargs)

    def input_single(self, mm: Magnets, value):
        super().input_single(mm, 1 if value else self.magnitude_ratio)


class PerpFieldInputter(FieldInputter):
    def __init__(self, datastream: BinaryDatastream, magnitude=1, angle=0, n=2, relax=True, frequency=1, **kwargs):
        """io.py:259-268: bit 0 at angle `angle`, bit 1 at `angle + pi/2`, each as a positive then
        a negative field; with relax=True the magnets are minimised (Magnets.minimize) under each
        of the two fields instead of stepped."""
        self.relax = relax
        super().__init__(datastream, magnitude=magnitude, angle=angle, n=n, frequency=frequency, **kwargs)

    def input_single(self, mm: Magnets, value):
        """io.py:270-288"""
        if (self.sine or self.frequency) and mm.params.UPDATE_SCHEME != Scheme.NEEL:
            raise AttributeError("Can not use temporal sine=True or nonzero frequency if UPDATE_SCHEME != Scheme.NEEL.")
        if not mm.in_plane:
            raise AttributeError("Can not use PerpFieldInputter on an out-of-plane ASI.")
        Zeeman = _zeeman(mm)
        angle = self.angle + value * math.pi / 2
        if self.relax:
            for mag in (self.magnitude, -self.magnitude):
                Zeeman.set_field(magnitude=mag, angle=angle)
                mm.minimize()
            return
        Zeeman.set_field(magnitude=self.magnitude, angle=angle)
        mm.progress(t_max=0.5 / self.frequency, MCsteps_max=self.n)
        Zeeman.set_field(magnitude=-self.magnitude, angle=angle)
        mm.progress(t_max=0.5 / self.frequency, MCsteps_max=self.n)


class ClockingFieldInputter(FieldInputter):
    def __init__(self, datastream: BinaryDatastream, magnitude=1, angle=0, spread=math.pi / 8, n=2, frequency=1):
        """examples/Clocking_IP_Pinwheel.py:28-36: a field at angle+spread for half a period, then at
        angle-spread, for bit 0; the same rotated by 180 degrees for bit 1, so each half addresses
        one sublattice."""
        self.spread = spread
        super().__init__(datastream, magnitude=magnitude, angle=angle, n=n, frequency=frequency)

    def bit_to_angles(self, bit):
        if not bit:
            return (self.angle + self.spread, self.angle - self.spread)
        return (self.angle + self.spread + math.pi, self.angle - self.spread + math.pi)

    def input_single(self, mm: Magnets, value):
        """examples/Clocking_IP_Pinwheel.py:43-55"""
        if self.frequency and mm.params.UPDATE_SCHEME != Scheme.NEEL:
            raise AttributeError("Can only use frequency if UPDATE_SCHEME is NEEL.")
        if not mm.in_plane:
            raise AttributeError("Can only use ClockingFieldInputter on in-plane ASI.")
        Zeeman = _zeeman(mm)
        angle1, angle2 = self.bit_to_angles(value)
        Zeeman.set_field(magnitude=self.magnitude, angle=angle1)
        mm.progress(t_max=0.5 / self.frequency, MCsteps_max=self.n)
        Zeeman.set_field(magnitude=self.magnitude, angle=angle2)
        mm.progress(t_max=0.5 / self.frequency, MCsteps_max=self.n)


class OOPSquareChessFieldInputter(Inputter):
    def __init__(self, datastream: Datastream, magnitude: float = 1, n=4, frequency=1):
        """io.py:385-396: a checkerboard field of magnitude magnitude*(2*value-1)."""
        super().__init__(datastream)
        self.magnitude = magnitude
        self.n = n
        self.frequency = frequency

    def input_single(self, mm: OOP_Square, value):
        """io.py:398-404"""
        Zeeman = _zeeman(mm)
        AFM_mask = (((mm.ixx + mm.iyy) % 2) * 2 - 1).astype(float)
        Zeeman.set_field(magnitude=AFM_mask * (2 * value - 1) * self.magnitude)
        mm.progress(t_max=1 / self.frequency, MCsteps_max=self.n)


class OOPSquareChessStepsInputter(Inputter):
    def __init__(self, datastream: Datastream, magnitude: float = 1., magnitude_range: float = 0.,
                 transition_range: float = 1., n: float = 4., frequency: float = 1.):
        """io.py:448-466: the checkerboard field in two steps, one half of the checkerboard each, with
        a piecewise linear value-to-magnitude map for scalar streams."""
        super().__init__(datastream)
        self.magnitude, self.magnitude_range, self.transition_range = magnitude, magnitude_range, transition_range
        self.n = n
        self.frequency = frequency
        self.transform = self._get_transform_func()

    def input_single(self, mm: OOP_Square, value):
        """io.py:468-481"""
        Zeeman = _zeeman(mm)
        magnitude = self.transform(value)
        AFM_mask = ((mm.ixx + mm.iyy) % 2).astype(float)
        Zeeman.set_field(magnitude=AFM_mask * magnitude)
        mm.progress(t_max=1 / self.frequency, MCsteps_max=self.n)
        Zeeman.set_field(magnitude=(AFM_mask - 1) * magnitude)
        mm.progress(t_max=1 / self.frequency, MCsteps_max=self.n)

    def _get_transform_func(self):
        if isinstance(self.datastream, (BinaryDatastream, IntegerDatastream)):
            return lambda bit: (2 * bit - 1) * self.magnitude
        if isinstance(self.datastream, ScalarDatastream):
            return self._val_to_mag_piecewiselinear

    def _val_to_mag_piecewiselinear(self, value: float):
        """io.py:489-500"""
        s = 0.5 - self.transition_range / 2
        if value < s:
            return -self.magnitude + value * self.magnitude_range / s
        if value > 1 - s:
            return self.magnitude + (value - 1) * self.magnitude_range / s
        return 2 * (self.magnitude - self.magnitude_range) / self.transition_range * (value - 0.5)


# ---- output readers (io.py:105-135, 311-445) -------------------------------------------------------
class OutputReader(ABC):
    def __init__(self, mm: Magnets = None):
        self.mm = mm
        if mm is not None:
            self.configure_for(mm)

    def configure_for(self, mm: Magnets):
        self.mm = mm

    @abstractmethod
    def read_state(self, mm: Magnets = None):
        """Reads the current state of `mm`; sets and returns self.state."""
        if (mm is not None) and (mm is not self.mm):
            self.configure_for(mm)

    @property
    def n(self) -> int:
        return self.state.size

    @property
    def node_coords(self) -> np.ndarray:
        return self._node_coords


class FullOutputReader(OutputReader):
    """io.py:311-331: the spins of every occupied cell."""

    def configure_for(self, mm: Magnets):
        self.mm = mm
        self.indices = np.nonzero(mm.occupation)
        self.state = np.asarray(mm.m)[self.indices]
        self._node_coords = np.asarray([mm.xx[self.indices], mm.yy[self.indices]]).T

    def read_state(self, mm: Magnets = None, m=None):
        super().read_state(mm)
        if m is None:
            m = self.mm.m
        self.state = np.asarray(m)[self.indices]
        return self.state


def _bins(nbins, ncells):
    """Bin index of every cell as the reference computes it, linspace(0, nbins*(1-1e-15), ncells)
    truncated (io.py:347-348), and the start of each bin (bins are contiguous and ascending)."""
    idx = np.linspace(0, nbins * (1 - 1e-15), ncells).astype(int)
    starts = np.searchsorted(idx, np.arange(nbins + 1), side='left')
    starts[-1] = ncells
    return idx, starts


class _BinnedReader(OutputReader):
    """Region bins of RegionalOutputReader / OOPSquareChessOutputReader (io.py:347-349)."""

    def _configure_bins(self, mm):
        ix, self._xs = _bins(self.nx, mm.nx)
        iy, self._ys = _bins(self.ny, mm.ny)
        self.region = ix[None, :] + self.nx * iy[:, None]


class RegionalOutputReader(_BinnedReader):
    def __init__(self, nx: int, ny: int, mm: Magnets = None):
        """io.py:335-342: the moment-weighted magnetisation averaged over nx x ny regions
        (x and y components for in-plane ASI)."""
        self.nx, self.ny = nx, ny
        self.n_regions = self.nx * self.ny
        super().__init__(mm)

    def configure_for(self, mm: Magnets):
        """io.py:344-357"""
        self.mm = mm
        self.state = np.zeros(self.n_regions * 2) if mm.in_plane else np.zeros(self.n_regions)
        self._configure_bins(mm)
        moment = np.asarray(mm.moment)
        self.normalization_factor = np.zeros(self.n)
        self._node_coords = np.zeros((self.n, 2))
        for i in range(self.n):
            regionindex = i % self.n_regions
            sel = self.region == regionindex
            self.normalization_factor[i] = np.sum(mm.occupation[sel] * moment[sel])
            self._node_coords[i, 0] = np.mean(mm.xx[sel])
            self._node_coords[i, 1] = np.mean(mm.yy[sel])
        w = [moment * mm.orientation[:, :, 0], moment * mm.orientation[:, :, 1]] if mm.in_plane else [moment]
        self._weights = mm._dev.bins(self._xs, self._ys, np.stack(w))

    def read_state(self, mm: Magnets = None, m=None) -> np.ndarray:
        """io.py:359-377 (the sums on the device; an explicit `m` is summed on the host)."""
        super().read_state(mm)
        if self.mm is None:
            raise AttributeError("OutputReader has not yet been initialized with a Magnets object.")
        if m is not None:
            m = np.asarray(m) * np.asarray(self.mm.moment)
            parts = [m * self.mm.orientation[:, :, 0], m * self.mm.orientation[:, :, 1]] if self.mm.in_plane else [m]
            sums = np.asarray([[np.sum(p[self.region == r]) for r in range(self.n_regions)] for p in parts])
        else:
            self.mm._sync_m()
            sums = self.mm._dev.binned_sums(self._weights)  # [component][region y][region x]
        self.state = sums.reshape(-1) / self.normalization_factor
        return self.state


class OOPSquareChessOutputReader(_BinnedReader):
    def __init__(self, nx: int, ny: int, mm: OOP_Square = None):
        """io.py:404-416: regional averages of the spins under the checkerboard (AFM) mask."""
        self.nx, self.ny = int(nx), int(ny)
        super().__init__(mm)

    def configure_for(self, mm: OOP_Square):
        """io.py:418-436"""
        if not isinstance(mm, (expected_type := OOP_Square)):
            if mm.in_plane:
                raise TypeError(f"{type(self).__name__} only works on out-of-plane ASI, but received {type(mm).__name__}.")
            warnings.warn(f"{type(self).__name__} expects {expected_type.__name__} ASI, but received {type(mm).__name__}.",
                          stacklevel=2)
        self.mm = mm
        self.state = np.zeros(self.nx * self.ny, dtype=float)
        self.AFM_mask = ((mm.ixx + mm.iyy) % 2) * 2 - 1
        self._configure_bins(mm)
        self.normalization_factor = np.zeros(self.n, dtype=int)
        self._node_coords = np.zeros((self.n, 2))
        for regionindex in range(self.n):
            sel = self.region == regionindex
            self.normalization_factor[regionindex] = np.sum(mm.occupation[sel])
            self._node_coords[regionindex, 0] = np.mean(mm.xx[sel])
            self._node_coords[regionindex, 1] = np.mean(mm.yy[sel])
        self._weights = mm._dev.bins(self._xs, self._ys, self.AFM_mask[None].astype(float))

    def read_state(self, mm: OOP_Square = None, m=None) -> np.ndarray:
        """io.py:438-445 (the sums on the device; an explicit `m` is summed on the host)."""
        super().read_state(mm)
        if m is not None:
            masked_m = self.AFM_mask * np.asarray(m)
            sums = np.asarray([np.sum(masked_m[self.region == r]) for r in range(self.n)])
        else:
            self.mm._sync_m()
            sums = self.mm._dev.binned_sums(self._weights).reshape(-1)
        self.state = sums / self.normalization_factor
        return self.state
