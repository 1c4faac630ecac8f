"""Magnets, SimParams and Scheme: the reference's stepping API over the HIP stepping kernels.

Mirrors hotspice/core.py (SimParams 29-53, Magnets 56-980): same constructor, fields, setters,
update()/progress() entry points and return values. The state (spins, combined local fields,
barriers, temperatures, moments) lives on the GPU; update() and progress() run entirely in
libhotspice_hip.so (hs_neel_run / hs_metropolis_run). Host arrays such as ``m`` are mirrors:
reading downloads, writing (or editing the returned array and calling update_energy())
uploads. There is no CPU fallback: without the library and a HIP device, Magnets raises.
"""
from __future__ import annotations

import math
import warnings
from abc import ABC, abstractmethod
from dataclasses import dataclass, field
from enum import Enum, auto
from typing import Literal

import numpy as np

from . import _lib
from ._device import DeviceState, kernel_window
from .energies import DiMonopolarEnergy, DipolarEnergy, Energy, ExchangeEnergy, ZeemanEnergy
from .utils import as_2D_array, check_repetition, full_obj_name, lower_than

kB = 1.380649e-23



class _Released:
    """Stands in for the device state of an instance whose device memory was released."""

    def __getattr__(self, name):
        raise _lib.HotspiceError("this Magnets instance released its device memory (release_device())")

class Scheme(Enum):
    """core.py:29-32"""
    NEEL = auto()
    METROPOLIS = auto()
    WOLFF = auto()


@dataclass(slots=True)
class SimParams:
    """core.py:34-53 (same fields, defaults and validation)."""
    SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF: int = 20
    REDUCED_KERNEL_SIZE: int = 20
    UPDATE_SCHEME: Scheme | str = Scheme.NEEL
    MULTISAMPLING_SCHEME: Literal['single', 'grid', 'Poisson', 'cluster'] = 'grid'
    ENERGY_BARRIER_METHOD: Literal['simple', 'parabolic'] = 'simple'
    METROPOLIS_TIME: bool = False

    def __post_init__(self):
        self.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF = int(self.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF)
        self.REDUCED_KERNEL_SIZE = int(self.REDUCED_KERNEL_SIZE)
        if not isinstance(self.UPDATE_SCHEME, Scheme):
            allowed = {'Néel': Scheme.NEEL, 'Neel': Scheme.NEEL, 'Metropolis': Scheme.METROPOLIS,
                       'Glauber': Scheme.METROPOLIS, 'Wolff': Scheme.WOLFF}
            if self.UPDATE_SCHEME not in allowed:
                raise ValueError(f"UPDATE_SCHEME='{self.UPDATE_SCHEME}' is invalid: allowed values are {allowed.keys()}.")
            self.UPDATE_SCHEME = allowed[self.UPDATE_SCHEME]
        if self.MULTISAMPLING_SCHEME not in (allowed := ['single', 'grid', 'Poisson', 'cluster']):
            raise ValueError(f"MULTISAMPLING_SCHEME='{self.MULTISAMPLING_SCHEME}' is invalid: allowed values are {allowed}.")


@dataclass(slots=True)
class Unitcell:
    x: int
    y: int


@dataclass
class History:
    """core.py:982-1000"""
    E: list = field(default_factory=list)
    T: list = field(default_factory=list)
    t: list = field(default_factory=list)
    m: list = field(default_factory=list)

    def entry(self, mm):
        self.E.append(mm.E_tot)
        self.T.append(mm.T_avg)
        self.t.append(mm.t if mm.t != 0 else mm.switches)
        self.m.append(mm.m_avg)

    def clear(self):
        self.E.clear()
        self.T.clear()
        self.t.clear()
        self.m.clear()


class Magnets(ABC):
    """Device-backed counterpart of hotspice.core.Magnets (core.py:56-980).

    Build-side additions (keyword-only, not in the reference):
      seed            Philox key of this instance (default: random); trajectories are a pure
                      function of (seed, state), so runs are reproducible.
      device          HIP device index (default 0).
      recompute_every Neel switches (or Metropolis steps) between FP64 recomputes of the
                      FP32 fields, which reset incremental drift (default 1024; None disables).
                      Such a periodic recompute refreshes the fields from an FP64 shadow of the
                      dipolar correlations, adding only the sources that changed since the last
                      one, while few did (shadow_refresh = False: always a full recompute).
      fp64_fields     hold the combined field in FP64 and step it with FP64 increments, as the
                      reference's float64 arrays do (default False: FP32 + Kahan compensation).
                      Metropolis stepping on grids of at most 80x80 cells, with DipolarEnergy,
                      ZeemanEnergy and ExchangeEnergy and no E_perp; other stepping raises.
    """

    shadow_refresh = True  # periodic recomputes may refresh from the FP64 shadow (per instance or class)

    def __init__(self, nx: int, ny: int, dx, dy, *, T=300, E_B=0., moment=None, Msat=800e3, V=2e-22,
                 pattern: str = None, energies=None, PBC: bool = False, m_perp_factor: float = None,
                 E_B_std: float = None, major_axis: float = 0, minor_axis: float = None, angle: float = 0.,
                 params: SimParams = None, in_plane: bool = False, seed: int = None, device: int = None,
                 recompute_every: int | None = 1024, fp64_fields: bool = False):
        self.seed = int(np.random.SeedSequence().entropy) & 0xFFFFFFFFFFFFFFFF if seed is None else int(seed) & 0xFFFFFFFFFFFFFFFF
        self.neel_persistent = True  # one persistent launch per progress() chunk (False: one launch per step)
        self.neel_blocks = 0         # workgroups in the instance's team (0: default, 'auto': autotuned)
        self.metro_blocks = 0        # Metropolis update grid (0: default, 'auto': autotuned per geometry)
        self._metro_tuned = {}       # selection geometry -> autotuned update grid
        self.rng = np.random.default_rng(self.seed)  # host draws at construction only ('random' pattern, E_B_std)
        self.params = SimParams() if params is None else params
        self.recompute_every = recompute_every
        self.fp64_fields = bool(fp64_fields)
        self._metro_rate = None  # measured Metropolis samples per step (sizes launches only)
        self._energies = list()
        self.in_plane = in_plane
        self.nx, self.ny = int(nx), int(ny)
        self.ixx, self.iyy = np.meshgrid(np.arange(0, self.nx), np.arange(0, self.ny))
        self._set_dx(dx)
        self._set_dy(dy)
        self.occupation = self._get_occupation().astype(bool).astype(int)
        with np.errstate(invalid='ignore', divide='ignore'):
            self._occupation_inf = 1 / self.occupation
        self.nonzero = np.nonzero(self.occupation)
        self._nonzero_array = np.asarray(self.nonzero).reshape(2, -1)
        self.n = int(np.sum(self.occupation))
        if self.n == 0:
            raise ValueError(f"Can not initialize {full_obj_name(self)} of size {self.nx}x{self.ny}, as this does not contain any spins.")
        if self.in_plane:
            self._initialize_ip(angle=angle)
        self.unitcell = self._get_unitcell()
        self._PBC = False
        self.PBC = PBC
        self._m_host = None
        self._m_ref = None
        self._occ_ref = None
        self._dev = DeviceState(self, device)
        self._m_perp_factor = 0.0
        self._calc_r_cache = {}
        self._since_recompute = 0
        self.recomputes = 0  # FP64 recomputes run so far (full, or refreshed from the shadow)
        self._sw_at_recompute = None  # switch count at the last Metropolis recompute (None: unknown)
        self._sel_key = None
        self.initialize_m(self._get_groundstate() if pattern is None else pattern, update_energy=False)
        self.T = T
        self.E_B = E_B
        if E_B_std is not None:
            self.E_B = self.E_B * self.rng.normal(1, E_B_std, size=self.E_B.shape)
        self.moment = Msat * V if moment is None else moment
        self.m_perp_factor = in_plane if m_perp_factor is None else m_perp_factor
        self.major_axis = major_axis
        self.minor_axis = major_axis if minor_axis is None else minor_axis
        self.history = History()
        self._dev.write_state(_lib.State())
        self._zeros = np.zeros_like(self.m)
        if energies is None:
            energies = (DipolarEnergy(),)
        elif isinstance(energies, Energy):
            energies = (energies,)
        else:
            energies = tuple(energies)
        self._upload_flags()
        self._device_params_changed()
        for energy in energies:
            self.add_energy(energy, _recompute=False)
        self._device_recompute()

    # ---- geometry ------------------------------------------------------------------------------
    def _set_dx(self, value):
        value = np.asarray(value, dtype=float).reshape(-1)
        self._dx = np.tile(value, math.ceil(self.nx / value.size))[:self.nx]
        self.x = np.zeros(self.nx)
        self.x[1:] = np.cumsum(self._dx)[:-1]
        self.xx, _ = np.meshgrid(self.x, np.empty(self.ny))
        self.x_min, self.x_max = float(self.x[0]), float(self.x[-1])

    def _set_dy(self, value):
        value = np.asarray(value, dtype=float).reshape(-1)
        self._dy = np.tile(value, math.ceil(self.ny / value.size))[:self.ny]
        self.y = np.zeros(self.ny)
        self.y[1:] = np.cumsum(self._dy)[:-1]
        _, self.yy = np.meshgrid(np.empty(self.nx), self.y)
        self.y_min, self.y_max = float(self.y[0]), float(self.y[-1])

    @property
    def dx(self):
        return self._dx

    @dx.setter
    def dx(self, value):
        self._set_dx(value)
        for e in self._energies:
            e.initialize(self)
        self._device_params_changed()
        self._device_recompute()

    @property
    def dy(self):
        return self._dy

    @dy.setter
    def dy(self, value):
        self._set_dy(value)
        for e in self._energies:
            e.initialize(self)
        self._device_params_changed()
        self._device_recompute()

    def _get_unitcell(self, max_cell=100):
        """core.py:146-161: smallest (ux, uy) in the same triangular search order."""
        dx, dy = self.dx.reshape(1, -1), self.dy.reshape(-1, 1)
        arr = self.orientation if self.in_plane else self.occupation
        for n in range(1, min(self.nx + self.ny + 1, max_cell + 1)):
            for uy in range(max(0, n - self.nx) + 1, min(n, self.ny) + 1):
                ux = n - uy + 1
                if check_repetition(arr, ux, uy) and check_repetition(dx, ux, 1) and check_repetition(dy, 1, uy):
                    return Unitcell(ux, uy)
        warnings.warn(f"Could not detect a reasonably sized unit cell. Defaulting to entire domain {self.nx}x{self.ny}.", stacklevel=2)
        return Unitcell(self.nx, self.ny)

    def _initialize_ip(self, angle: float = 0.):
        """core.py:182-193"""
        self.original_angles = self._get_angles() + angle
        self.angles = self.original_angles * self.occupation
        self.orientation = np.zeros(self.ixx.shape + (2,))
        self.orientation[:, :, 0] = np.cos(self.angles) * self.occupation
        self.orientation[:, :, 1] = np.sin(self.angles) * self.occupation
        self._obs_all = None  # device readout weights follow the orientation

    @property
    def PBC(self):
        return self._PBC

    @PBC.setter
    def PBC(self, value: bool):
        """core.py:361-375"""
        if value and (self.nx % self.unitcell.x != 0 or self.ny % self.unitcell.y != 0):
            warnings.warn(f"PBC can not be enabled, because there is a non-integer number of unit cells along the x- "
                          f"and/or y-axis (unit cell {self.unitcell.x}x{self.unitcell.y}).", stacklevel=2)
            self._PBC = False
            return
        self._PBC = bool(value)
        if hasattr(self, "_dev"):
            for e in self._energies:
                e.initialize(self)
            self._device_params_changed()
            self._device_recompute()

    @property
    def shape(self):
        return (self.ny, self.nx)

    # ---- spins ---------------------------------------------------------------------------------
    def _get_m_uniform(self, angle=0):
        """core.py:138-144"""
        angle += 1e-6
        if self.in_plane:
            return self.occupation * (2 * ((self.orientation[:, :, 0] * math.cos(angle) + self.orientation[:, :, 1] * math.sin(angle)) >= 0) - 1)
        return np.ones_like(self.xx) * ((math.floor(angle / math.pi - .5) % 2) * 2 - 1)

    def initialize_m(self, pattern='random', *, angle: float = 0, update_energy: bool = True):
        """core.py:163-180"""
        if isinstance(pattern, str):
            self._set_m(pattern)
            m = self._m_pattern
        else:
            try:
                m = as_2D_array(pattern, self.shape)
            except (ValueError, TypeError):
                raise ValueError("Argument <pattern> could not be parsed as a valid array for <self.m>.")
        m = np.sign(np.asarray(m, dtype=float))
        m = np.multiply(m, self.occupation)
        m[np.where(self._get_m_uniform() != self._get_m_uniform(angle))] *= -1
        self.m = m
        if update_energy:
            self.update_energy()

    def _set_m(self, pattern: str):
        """core.py:924-952 (subclasses add lattice patterns); stores the result in self._m_pattern."""
        match str(pattern).strip().lower():
            case 'uniform':
                self._m_pattern = self._get_m_uniform()
            case 'vortex':
                m = np.ones_like(self.xx)
                if self.in_plane:
                    diff = (self.xx - (self.x_max - self.x_min) / 2) + 1j * (self.yy - (self.y_max - self.y_min) / 2)
                    desired = diff * (-1j + 1e-6)
                    dot = desired.real * self.orientation[:, :, 0] + desired.imag * self.orientation[:, :, 1]
                    dot[dot == 0] = 1
                    occ = np.where(self.occupation == 1)
                    m[occ] = np.sign(dot)[occ]
                else:
                    m[np.where(self.xx >= (self.x_max - self.x_min) / 2)] *= -1
                self._m_pattern = m
            case str(unknown):
                self._m_pattern = self.rng.integers(0, 2, size=self.xx.shape) * 2 - 1
                if unknown != 'random':
                    warnings.warn(f"Pattern '{unknown}'' not recognized, defaulting to 'random'.", stacklevel=2)

    @property
    def m(self) -> np.ndarray:
        """(ny, nx) float64 host mirror of the device spins (-1/0/+1)."""
        if self._m_host is None:
            self._m_host = self._dev.download(self._dev.m).reshape(self.ny, self.nx).astype(float)
            self._m_ref = self._m_host.copy()
        return self._m_host

    @m.setter
    def m(self, value):
        self._m_host = np.array(as_2D_array(value, self.shape), dtype=float)
        self._m_ref = None
        self._sync_m()

    def _sync_m(self):
        """Uploads the host mirror if it was assigned or edited in place since the last sync."""
        if self._m_host is None:
            return
        if self._m_ref is not None and np.array_equal(self._m_host, self._m_ref):
            return
        self._dev.upload(self._dev.m, np.sign(self._m_host).astype(np.int8))
        self._m_ref = self._m_host.copy()

    def _after_device_step(self):
        self._m_host = None
        self._m_ref = None

    def _upload_flags(self):
        occ = np.asarray(self.occupation).astype(bool)
        neel = np.isfinite(self._occupation_inf)
        flags = occ.astype(np.uint8) * _lib.HS_FLAG_OCC + neel.astype(np.uint8) * _lib.HS_FLAG_NEEL
        self._dev.upload(self._dev.flags, flags)
        self._dev.invalidate_shadow()  # targets that were empty have no shadow entry
        occ_list = np.flatnonzero(np.isfinite(self._occupation_inf)).astype(np.int32)
        self._dev.upload(self._dev.occ_list[:occ_list.size], occ_list)
        self._occ_ref = np.array(self.occupation, copy=True)

    def _sync_occupation(self):
        if self._occ_ref is None or not np.array_equal(self._occ_ref, self.occupation):
            self._upload_flags()
            self._sel_key = None

    # ---- fields --------------------------------------------------------------------------------
    @property
    def T(self):
        return self._T

    @T.setter
    def T(self, value):
        """core.py:331-334"""
        self._T = as_2D_array(value, self.shape)
        self.kBT = kB * self._T
        with np.errstate(divide='ignore'):
            self.beta = 1 / self.kBT
        self._T_any_zero = bool(np.any(self._T == 0))
        self._dev.upload(self._dev.beta, self.beta.astype(np.float32))
        self._uniform_beta = bool(np.all(self.beta == self.beta.reshape(-1)[0]))
        if self._uniform_beta:
            self._dev.beta64c = None
        else:
            if self._dev.beta64c is None:
                self._dev.beta64c = self._dev.torch.zeros(self.nx * self.ny, dtype=self._dev.torch.float64,
                                                          device=self._dev.device)
            self._dev.upload(self._dev.beta64c, self.beta.astype(np.float64))
        if hasattr(self, "_energies"):
            self._device_params_changed()

    @property
    def E_B(self):
        self._EB_exposed = True  # the caller may edit the array in place: compare before the next step
        return self._E_B

    @E_B.setter
    def E_B(self, value):
        """core.py:340-343"""
        self._E_B = np.array(as_2D_array(value, self.shape), dtype=float)
        with np.errstate(divide='ignore'):
            self._E_B_inv_over_four = .25 / self._E_B
        self._dev.upload(self._dev.EB, self._E_B.astype(np.float32))
        self._EB_ref = self._E_B.copy()
        self._EB_exposed = False  # a new array: no reference to it has been handed out yet

    def _sync_EB(self):
        """Uploads E_B if it was edited in place since the last upload (only possible once the
        array has been handed out by the E_B getter)."""
        if self._EB_exposed and not np.array_equal(self._E_B, self._EB_ref):
            self.E_B = self._E_B

    @property
    def moment(self):
        return self._moment

    @moment.setter
    def moment(self, value):
        """core.py:349-351"""
        self._moment = np.array(as_2D_array(value, self.shape), dtype=float)
        self._momentSq = self._moment * self._moment
        self._dev.upload(self._dev.mu64, self._moment)
        self._sync_pair_weights()

    @property
    def USE_PERP_ENERGY(self):
        return bool(self.m_perp_factor) and self.in_plane

    @property
    def m_perp_factor(self):
        return self._m_perp_factor

    @m_perp_factor.setter
    def m_perp_factor(self, value):
        """core.py:279-287"""
        if not isinstance(value, (int, float, bool)):
            raise ValueError("m_perp_factor must be scalar. It is not currently supported as a 2D array.")
        if value and not self.in_plane:
            warnings.warn("You set a nonzero `m_perp_energy`, but this has no effect for an OOP ASI.")
        used = self.USE_PERP_ENERGY
        self._m_perp_factor = float(value)
        if hasattr(self, "_energies") and self._energies:
            self._device_params_changed()
            if not used and self.USE_PERP_ENERGY:
                self._device_recompute()

    def release_device(self):
        """Frees this instance's device arrays (a server retiring a finished job). Reading host
        attributes still works; stepping or reading device state afterwards raises."""
        self._m_host = self.m.copy()
        self._m_ref = self._m_host.copy()
        self._dev = _Released()
        self._sel, self._sel_tensors, self._sel_key = None, (), None

    # ---- counters (device run state) ------------------------------------------------------------
    def _state(self):
        return self._dev.read_state()

    @property
    def t(self):
        return float(self._state().t)

    @t.setter
    def t(self, value):
        st = self._state()
        st.t = float(value)
        self._dev.write_state(st)

    @property
    def switches(self):
        return int(self._state().switches)

    @switches.setter
    def switches(self, value):
        st = self._state()
        st.switches = int(value)
        self._dev.write_state(st)

    @property
    def attempted_switches(self):
        return int(self._state().attempts)

    @attempted_switches.setter
    def attempted_switches(self, value):
        st = self._state()
        st.attempts = int(value)
        self._dev.write_state(st)

    @property
    def MCsteps(self):
        return self.attempted_switches / self.n

    def reset_stats(self):
        st = self._state()
        st.t, st.switches, st.attempts = 0.0, 0, 0
        self._dev.write_state(st)

    # ---- energies ------------------------------------------------------------------------------
    def add_energy(self, energy: Energy, exist_ok=False, verbose=True, _recompute=True):
        """core.py:195-209"""
        if not isinstance(energy, (ZeemanEnergy, DipolarEnergy, ExchangeEnergy)):
            raise NotImplementedError(f"{type(energy).__name__} has no device implementation; supported energies are "
                                      "ZeemanEnergy, DipolarEnergy, DiMonopolarEnergy and ExchangeEnergy.")
        if isinstance(energy, DipolarEnergy):
            other = self._pair_energy()
            if other is not None and type(other) is not type(energy):
                raise NotImplementedError(f"{type(energy).__name__} and {type(other).__name__} cannot be combined on the "
                                          f"device (one pair-interaction kernel set): remove the {type(other).__name__} first.")
        for i, e in enumerate(self._energies):
            if type(e) is type(energy):
                if exist_ok:
                    return e
                if verbose:
                    warnings.warn(f"An instance of {type(energy).__name__} was already included in the simulation, and has now been overwritten.", stacklevel=2)
                energy.initialize(self)
                self._energies[i] = energy
                self._sync_pair_weights()
                self._device_params_changed()
                if _recompute:
                    self._device_recompute()
                return
        energy.initialize(self)
        self._energies.append(energy)
        self._sync_pair_weights()
        self._device_params_changed()
        if _recompute:
            self._device_recompute()
        return energy

    def remove_energy(self, name: str, verbose=True):
        """core.py:211-221"""
        name = name.lower().replace('energy', '')
        for i, e in enumerate(self._energies):
            if name == e.shortname:
                self._energies.pop(i)
                if isinstance(e, DipolarEnergy):
                    self._dev.drop_kernels()
                    self._sync_pair_weights()
                self._device_params_changed()
                self._device_recompute()
                return
        if verbose:
            warnings.warn(f"There is no '{name}' energy associated with this Magnets object. Valid energies are: {[e.shortname for e in self._energies]}", stacklevel=2)

    def has_energy(self, name: str):
        name = name.lower().replace('energy', '')
        return any(name == e.shortname for e in self._energies)

    def get_energy(self, name: str, verbose=True):
        """core.py:229-239"""
        name = name.lower().replace('energy', '')
        for e in self._energies:
            if name == e.shortname:
                return e
        if verbose:
            warnings.warn(f"There is no '{name}' energy associated with this Magnets object. Valid energies are: {[e.shortname for e in self._energies]}", stacklevel=2)
        return None

    def update_energy(self, index=None):
        """core.py:241-253. Host edits of m / occupation are uploaded, then the combined device
        field is recomputed in FP64 (index=None) -- also for an explicit index, where the
        recompute equals the reference's incremental update up to rounding."""
        self._sync_occupation()
        self._sync_m()
        self._sync_EB()
        self._device_recompute()

    def switch_energy(self, indices2D=None):
        """core.py:255-260: -2*sum(E) at the given indices (or everywhere)."""
        E = self.E
        return -2 * E if indices2D is None else -2 * E[indices2D[0], indices2D[1]]

    def perp_energy(self, indices2D=None):
        """core.py:262-271"""
        if not self.USE_PERP_ENERGY:
            return self._zeros if indices2D is None else np.zeros(np.asarray(indices2D).shape[1])
        Ep = self.E_perp
        return Ep if indices2D is None else Ep[indices2D[0], indices2D[1]]

    @property
    def E(self) -> np.ndarray:
        """Sum of the energies' E (core.py:305-307), FP64 from the device's compensated field."""
        return self._dev.components()[4]

    @property
    def E_perp(self) -> np.ndarray:
        """core.py:309-312"""
        return self._dev.components()[5] * self.m_perp_factor

    @property
    def E_tot(self) -> float:
        """core.py:426-428"""
        return float(self._dev.energy_total()) if self._energies else 0.0

    def _energy_component(self, energy, perp=False):
        comp = self._dev.components()
        if isinstance(energy, ZeemanEnergy):
            return comp[1] if perp else comp[0]
        if isinstance(energy, ExchangeEnergy):
            return comp[3] if perp else comp[2]
        total = comp[5] if perp else comp[4]
        zee = comp[1] if perp else comp[0]
        exc = comp[3] if perp else comp[2]
        return total - zee - exc

    # ---- device instance -------------------------------------------------------------------------
    def _pair_energy(self):
        """The DipolarEnergy or DiMonopolarEnergy of this instance (the device's pair kernels), or None."""
        for e in getattr(self, "_energies", ()):
            if isinstance(e, DipolarEnergy):
                return e
        return None

    def _sync_pair_weights(self):
        """Uploads each magnet's pair weight: its moment for DipolarEnergy, its squared charge
        (moment/d)^2 for DiMonopolarEnergy (energies.py:413-415, fixed when that energy is
        initialised)."""
        d = self._dev
        d.invalidate_shadow()  # the shadow holds correlations of the old weights
        pair = self._pair_energy()
        if isinstance(pair, DiMonopolarEnergy) and hasattr(pair, "charge_sq"):
            q2 = np.ascontiguousarray(pair.charge_sq, dtype=np.float64)
            if d.pw64 is None:
                d.pw64 = d.torch.empty(self.nx * self.ny, dtype=d.torch.float64, device=d.device)
                d.src_ones = d.torch.ones(self.nx * self.ny, dtype=d.torch.float32, device=d.device)
            d.upload(d.pw64, q2)
            d.upload(d.mu, q2.astype(np.float32))
        else:
            d.pw64 = None
            d.src_ones = None
            d.upload(d.mu, self._moment.astype(np.float32))

    def _device_build_kernels(self, dip: DipolarEnergy):
        """Host part of DipolarEnergy._initialize (energies.py:183-214): offsets of the tiled
        unit-cell positions around each sublattice's centre magnet; the device builds the kernels."""
        if isinstance(dip, DiMonopolarEnergy):
            return self._device_build_monopole_kernels(dip)
        ux, uy, nx, ny = self.unitcell.x, self.unitcell.y, self.nx, self.ny
        ncx = 2 * math.ceil(nx / ux) + 1
        ncy = 2 * math.ceil(ny / uy) + 1
        uc_dx = np.tile(self.dx[:ux], ncx)
        uc_dy = np.tile(self.dy[:uy], ncy)
        g = _lib.KernelGeom()
        sites = []
        for y in range(uy):
            for x in range(ux):
                if self.in_plane:
                    if self.orientation[y, x, 0] == 0 and self.orientation[y, x, 1] == 0:
                        continue
                elif self.occupation[y, x] == 0:
                    continue
                sites.append((y, x))
        if not sites:
            raise ValueError("No occupied site in the unit cell: no dipolar kernel to build.")
        if len(sites) > _lib.HS_MAX_UC:
            raise ValueError("Too many unit-cell sites for the device kernels.")
        rrx = np.zeros((len(sites), 2 * nx - 1))
        rry = np.zeros((len(sites), 2 * ny - 1))
        for n, (y, x) in enumerate(sites):
            sx = (ux - ((nx - 1) % ux) + x) % ux
            sy = (uy - ((ny - 1) % uy) + y) % uy
            xx = np.cumsum(uc_dx[sx:sx + 2 * nx - 1])
            yy = np.cumsum(uc_dy[sy:sy + 2 * ny - 1])
            rrx[n] = xx - xx[nx - 1]
            rry[n] = yy - yy[ny - 1]
            g.site_x[n], g.site_y[n], g.shift_x[n], g.shift_y[n] = x, y, sx, sy
        for s in range(uy * ux):
            y, x = divmod(s, ux)
            if self.in_plane:
                g.uc_ox[s] = float(self.orientation[y, x, 0])
                g.uc_oy[s] = float(self.orientation[y, x, 1])
            else:
                g.uc_o[s] = float(self.occupation[y, x])
        kwx, kwy = kernel_window(rrx, rry, nx, ny, float(dip.cutoff_range), self.PBC)
        g.I = (self.major_axis ** 2 + self.minor_axis ** 2) / 16
        g.cutoff = float(dip.cutoff_range)
        g.decay = float(dip.decay_exponent)
        g.nx, g.ny, g.ux, g.uy, g.n_sub, g.kwx, g.kwy = nx, ny, ux, uy, len(sites), kwx, kwy
        g.pbc, g.in_plane = int(self.PBC), int(self.in_plane)
        self._dev.build_kernels(g, rrx, rry, self.in_plane)
        self._kernel_sites = sites

    def _device_build_monopole_kernels(self, dim: DiMonopolarEnergy):
        """Host part of DiMonopolarEnergy._initialize (energies.py:416-463): positions of the
        enlarged grid's columns and rows seen from each occupied unit-cell site, and the dumbbell
        geometry of every site; the device builds the kernels and the pair weights follow."""
        ux, uy, nx, ny = self.unitcell.x, self.unitcell.y, self.nx, self.ny
        sites = [(y, x) for y in range(uy) for x in range(ux) if self.occupation[y, x] != 0]
        if not sites:
            raise ValueError("No occupied site in the unit cell: no DiMonopolar kernel to build.")
        if len(sites) > _lib.HS_MAX_UC:
            raise ValueError("Too many unit-cell sites for the device kernels.")

        def large(spacing, n, u):  # positions of indices -(n-1) .. n+u-1, index n-1 at 0 (energies.py:416-428)
            idx = np.arange(-(n - 1), n + u)
            d = np.asarray(spacing)[idx % n]
            pos = np.zeros(idx.size)
            pos[1:] = np.cumsum(d)[:-1]
            return pos - pos[n - 1]

        x_large, y_large = large(self.dx, nx, ux), large(self.dy, ny, uy)
        g = _lib.KernelGeom()
        rrx = np.zeros((len(sites), 2 * nx - 1))
        rry = np.zeros((len(sites), 2 * ny - 1))
        for n, (y, x) in enumerate(sites):
            rrx[n] = x_large[x:x + 2 * nx - 1]
            rry[n] = y_large[y:y + 2 * ny - 1]
            g.site_x[n], g.site_y[n] = x, y
            g.shift_x[n] = (ux - ((nx - 1) % ux) + x) % ux
            g.shift_y[n] = (uy - ((ny - 1) % uy) + y) % uy
        ang = np.asarray(self.original_angles[:uy, :ux], dtype=float)
        ca, sa = np.cos(ang), np.sin(ang)
        cp, sp = np.cos(ang + np.pi / 2), np.sin(ang + np.pi / 2)
        mg = _lib.MonopoleGeom()
        for s in range(uy * ux):
            y, x = divmod(s, ux)
            g.uc_o[s] = float(self.occupation[y, x])
            mg.ca[s], mg.sa[s], mg.cp[s], mg.sp[s] = ca[y, x], sa[y, x], cp[y, x], sp[y, x]
            mg.h[s] = float(dim.dist_unitcell[y, x]) / 2
        mg.h_small = float(dim.small_d) / 2
        g.nx, g.ny, g.ux, g.uy, g.n_sub, g.kwx, g.kwy = nx, ny, ux, uy, len(sites), nx - 1, ny - 1
        g.pbc, g.in_plane = int(self.PBC), 1
        self._dev.build_kernels(g, rrx, rry, True, mono=mg)
        self._kernel_sites = sites
        self._sync_pair_weights()

    def _device_params_changed(self):
        """(Re)fills the hs_instance_t the kernels read."""
        d = self._dev
        d.gen += 1  # sizes derived from the instance (recompute scratch) are asked for again
        I = d.inst
        P = d.ptr
        I.m, I.flags, I.E, I.Ec = P(d.m), P(d.flags), P(d.E), P(d.Ec)
        I.Ep, I.Epc, I.EB, I.beta, I.mu, I.mu64 = P(d.Ep), P(d.Epc), P(d.EB), P(d.beta), P(d.mu), P(d.mu64)
        I.occ_list = P(d.occ_list)
        I.beta64c = P(d.beta64c)
        I.E64 = P(d.E64)
        I.seed = self.seed
        I.nx, I.ny, I.ux, I.uy = self.nx, self.ny, self.unitcell.x, self.unitcell.y
        I.pbc, I.in_plane, I.use_perp = int(self.PBC), int(self.in_plane), int(self.USE_PERP_ENERGY)
        I.m_perp_factor = float(self.m_perp_factor)
        I.barrier = 0 if self.params.ENERGY_BARRIER_METHOD == 'simple' else 1
        I.n_occ = int(np.count_nonzero(np.isfinite(self._occupation_inf)))
        I.uniform_beta = int(getattr(self, "_uniform_beta", True))
        I.beta64 = float(self.beta.reshape(-1)[0]) if I.uniform_beta else 0.0
        for s in range(_lib.HS_MAX_UC):
            I.sub_of[s] = -1
            I.ox_uc[s] = 0.0
            I.oy_uc[s] = 0.0
        for s in range(self.unitcell.x * self.unitcell.y):
            y, x = divmod(s, self.unitcell.x)
            if self.in_plane:
                I.ox_uc[s] = float(self.orientation[y, x, 0])
                I.oy_uc[s] = float(self.orientation[y, x, 1])
        dip = self._pair_energy()
        I.pw64 = P(d.pw64) if d.pw64 is not None else P(d.mu64)
        I.mu_src = P(d.src_ones) if d.src_ones is not None else P(d.mu)
        I.pair_src_unit = int(isinstance(dip, DiMonopolarEnergy))
        if dip is not None and d.kern is not None:
            k = d.kern
            I.dip_pf = float(dip.prefactor)
            I.K32, I.K64 = P(k["K32"]), P(k["K64"])
            I.Kpo32, I.Kps64 = P(k["Kpo32"]), P(k["Kps64"])
            I.kwx, I.kwy, I.n_sub = k["kwx"], k["kwy"], k["n_sub"]
            for n, (y, x) in enumerate(self._kernel_sites):
                I.sub_of[y * self.unitcell.x + x] = n
        else:
            I.dip_pf = 0.0
            I.K32 = I.K64 = I.Kpo32 = I.Kps64 = None
            I.kwx = I.kwy = I.n_sub = 0
        I.C64, I.m_ref = d.shadow_ptrs(self.in_plane, enabled=bool(I.dip_pf))
        ex = self.get_energy('exchange', verbose=False)
        if ex is not None and ex.J != 0:
            nn = np.asarray(ex.local_interaction, dtype=float)
            if nn.shape[0] % 2 == 0 or nn.shape[1] % 2 == 0 or nn.size > _lib.HS_MAX_NN:
                raise ValueError(f"Exchange mask of shape {nn.shape} is not supported (odd sides, <= {_lib.HS_MAX_NN} entries).")
            I.J = float(ex.J)
            I.nn_h, I.nn_w = nn.shape
            for q in range(_lib.HS_MAX_NN):
                I.nn_mask[q] = float(nn.reshape(-1)[q]) if q < nn.size else 0.0
        else:
            I.J = 0.0
            I.nn_h = I.nn_w = 0
        z = self.get_energy('zeeman', verbose=False)
        I.has_zeeman = int(z is not None)
        d.zee_bx = d.zee_by = None
        I.zee_bx = I.zee_by = None
        I.zee_Bx = I.zee_By = 0.0
        if z is not None:
            bx, by, uniform = z.field_components()
            if uniform:
                I.zee_Bx, I.zee_By = float(bx), float(by)
            else:
                t = d.torch
                d.zee_bx = t.from_numpy(np.ascontiguousarray(bx, dtype=np.float32).reshape(-1)).to(d.device)
                d.zee_by = t.from_numpy(np.ascontiguousarray(by, dtype=np.float32).reshape(-1)).to(d.device)
                I.zee_bx, I.zee_by = P(d.zee_bx), P(d.zee_by)

    def _device_recompute(self, periodic=False, st=None):
        """FP64 recompute of the combined fields (Energy.update for every energy): a full one, or,
        for a periodic recompute (`st`: the current run state, if read already), a refresh from
        the FP64 shadow when few spins changed since the last recompute."""
        if not hasattr(self, "_dev"):
            return
        self._dev.recompute(3 if self.USE_PERP_ENERGY else 1, self._changed_bound(st) if periodic else None)
        self._recompute_done(st)

    def _changed_bound(self, st=None, ahead=0):
        """Upper bound on the cells whose spin changed since the last FP64 recompute: the switches
        since then, plus `ahead` more in launches queued before the recompute; None when unknown.
        It only chooses between a refresh from the shadow and a full recompute (both give the
        fields of the current spins), so a host write to the spins or counters cannot make it wrong."""
        if not self.shadow_refresh or self._dev.C64 is None:
            return None
        if self.params.UPDATE_SCHEME == Scheme.NEEL:
            return self._since_recompute + ahead  # one switch per Neel step
        if self._sw_at_recompute is None:
            return None
        sw = int((self._state() if st is None else st).switches)
        return sw - self._sw_at_recompute + ahead if sw >= self._sw_at_recompute else None

    def _recompute_done(self, st=None):
        """Counters after a recompute (`st`: the run state it ran at, if read already)."""
        self._since_recompute = 0
        self.recomputes += 1
        if self.shadow_refresh and self._dev.C64 is not None and self.params.UPDATE_SCHEME != Scheme.NEEL:
            self._sw_at_recompute = int((self._state() if st is None else st).switches)

    def _device_zeeman_changed(self, z: ZeemanEnergy, old):
        """ZeemanEnergy.set_field: swap the Zeeman term of the combined device field."""
        if z not in self._energies:
            return
        bx0, by0, uniform0 = old
        keep = None
        if uniform0:
            args = (bx0, by0, None, None)
        else:
            t = self._dev.torch
            keep = (t.from_numpy(np.ascontiguousarray(bx0, dtype=np.float32).reshape(-1)).to(self._dev.device),
                    t.from_numpy(np.ascontiguousarray(by0, dtype=np.float32).reshape(-1)).to(self._dev.device))
            args = (0.0, 0.0) + keep
        self._device_params_changed()
        self._dev.zeeman_swap(*args)
        if keep is not None:
            self._dev.torch.cuda.current_stream(self._dev.device).synchronize()

    # ---- stepping ------------------------------------------------------------------------------
    def _pre_step(self):
        self._sync_m()
        self._sync_EB()

    def update(self, *args, **kwargs):
        """core.py:604-622: one step of the configured scheme."""
        if self._T_any_zero:
            warnings.warn("T=0 somewhere in the domain, so no switch will be simulated to prevent DIV/0 errors.", stacklevel=2)
            return None
        match self.params.UPDATE_SCHEME:
            case Scheme.NEEL:
                return self._update_Néel(*args, **kwargs)
            case Scheme.METROPOLIS:
                return self._update_Metropolis(*args, **kwargs)
            case Scheme.WOLFF:
                return self._update_Wolff(*args, **kwargs)
            case other:
                raise ValueError(f"The Magnets.params object has an invalid value for UPDATE_SCHEME='{other}'.")

    def _neel_team_ok(self):
        """The persistent team kernel steps uniform temperatures; a temperature that varies over
        the grid takes the per-step launches (two per step: the grid's minimum barrier reorders
        the waiting times, core.py:686-699)."""
        return self.neel_persistent and self._uniform_beta

    def _neel_nblocks(self):
        """Team size of the next Neel launch ('auto' runs the autotuner once)."""
        if self.neel_blocks == 'auto':
            from .tune import tune_neel_blocks
            tune_neel_blocks(self)
        return self.neel_blocks

    def _metro_nblocks(self, r):
        """Update grid (hs_run_t.nblocks) of the next Metropolis launch at sampling distance r
        ('auto' runs the autotuner once per selection geometry)."""
        if self.metro_blocks == 'auto':
            self._select_geometry(r)
            if self._sel_key not in self._metro_tuned:
                from .tune import tune_metro_blocks
                tune_metro_blocks(self, r)
        return self._metro_grid()

    def _metro_grid(self):
        """Update grid for the current selection geometry (0, the library's default, while an
        'auto' instance has not been tuned for it)."""
        if self.metro_blocks == 'auto':
            return self._metro_tuned.get(self._sel_key, 0)
        return int(self.metro_blocks or 0)

    def _update_Néel(self, t_max=1, attempt_freq=1e10):
        """core.py:679-709: returns the (y, x) divmod tuple of the switched magnet, or None."""
        self._pre_step()
        self._maybe_recompute()
        run = _lib.Run(t_max=float(t_max), mc_max=math.inf, attempt_freq=float(attempt_freq), max_steps=1, log=1,
                       nblocks=0, begin=1, first_free=1)
        self._dev.neel_run(run, persistent=self._neel_team_ok(), nblocks=self._neel_nblocks())
        self._after_device_step()
        idx = int(self._dev.log_idx[0].item())
        st = self._state()
        if st.err:
            raise _lib.HotspiceError(f"device Neel step failed (err={st.err})")
        if idx < 0:
            return None
        self._since_recompute += 1
        return divmod(np.int64(idx), np.int64(self.nx))

    def _explicit_selection(self, idx):
        """The selection of update(idx=...) (core.py:660-661): the caller's (2, k) array of [y; x]
        sample positions, in its order, as a device list (None when k = 0). The samples must be
        distinct magnets of the grid."""
        idx = np.asarray(idx)
        if idx.ndim != 2 or idx.shape[0] != 2 or not np.issubdtype(idx.dtype, np.integer):
            raise ValueError("idx must be an integer array of shape (2, k) holding [y; x] positions")
        if idx.shape[1] == 0:
            return None
        y, x = idx[0].astype(np.int64), idx[1].astype(np.int64)
        if np.any((y < 0) | (y >= self.ny) | (x < 0) | (x >= self.nx)):
            raise ValueError("idx holds a position outside the grid")
        if not np.all(np.asarray(self.occupation)[y, x] != 0):
            raise ValueError("idx holds a position without a magnet")
        flat = y * self.nx + x
        if np.unique(flat).size != flat.size:
            raise ValueError("idx holds a position more than once")
        sel = _lib.Select()
        sel.scheme = 3
        sel.nsc = int(flat.size)
        sel.cutoff = self.params.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF
        sel.reduced_k = self.params.REDUCED_KERNEL_SIZE
        sel.metropolis_time = int(bool(self.params.METROPOLIS_TIME))
        t = self._dev.torch
        cells = t.from_numpy(flat.astype(np.int32)).to(self._dev.device)
        sel.pat_cells = cells.data_ptr()
        sel._keep = cells  # the device list lives as long as the selection
        return sel

    def _skip_step(self):
        """A step with an empty sample list: nothing is attempted, but the step counter (the key
        of the random streams) advances as for any update()."""
        st = self._state()
        st.step += 1
        self._dev.write_state(st)

    def _update_Metropolis(self, idx=None, Q=0.05, r=0, attempt_freq=1e10):
        """core.py:659-677: returns idx_switch, a (2, a) array of [y; x]."""
        self._pre_step()
        if idx is not None:
            sel = self._explicit_selection(idx)
            if sel is None:
                self._skip_step()
                return np.zeros((2, 0), dtype=np.int64)
            nblocks = self._metro_grid() if self.metro_blocks != 'auto' else 0
        else:
            r = self.calc_r(Q) if r == 0 else r
            nblocks = self._metro_nblocks(r)
            sel = self._select_geometry(r)
        cap = max(1, sel.nsc)
        t = self._dev.torch
        cells = t.zeros(cap, dtype=t.int32, device=self._dev.device)
        count = t.zeros(1, dtype=t.int32, device=self._dev.device)
        run = _lib.Run(t_max=math.inf, mc_max=math.inf, attempt_freq=float(attempt_freq), max_steps=1, log=1,
                       nblocks=nblocks, begin=1, first_free=1)
        self._dev.metropolis_run(sel, run, cells, count)
        self._after_device_step()
        a = int(count.item())
        st = self._state()
        if st.err:
            raise _lib.HotspiceError(f"device Metropolis step failed (err={st.err})")
        c = cells[:a].cpu().numpy().astype(np.int64)
        # the periodic FP64 recompute counts update() steps as progress() counts its steps
        self._since_recompute += 1
        self._maybe_recompute(st)
        return np.asarray([c // self.nx, c % self.nx])

    def _wolff_check(self):
        """core.py:577-590: clusters need an exchange energy and no DipolarEnergy."""
        if self.get_energy('dipolar', verbose=False) is not None:
            raise NotImplementedError("Can not (yet?) select Wolff cluster if long-range interactions are present.")
        ex = self.get_energy('exchange', verbose=False)
        if ex is None:
            raise NotImplementedError("Can not select Wolff cluster if there is no exchange energy in the system.")

    def _update_Wolff(self):
        """core.py:711-718: grows a Wolff cluster from a random magnet (_select_cluster_exchange,
        core.py:588-602) and switches it; returns the cluster as a (2, k) array of [y; x]."""
        self._wolff_check()
        self._pre_step()
        t = self._dev.torch
        cells = t.zeros(self.nx * self.ny, dtype=t.int32, device=self._dev.device)
        count = t.zeros(1, dtype=t.int32, device=self._dev.device)
        run = _lib.Run(t_max=math.inf, mc_max=math.inf, attempt_freq=1e10, max_steps=1, log=1, nblocks=0, begin=1,
                       first_free=1)
        p = self.params
        self._dev.wolff_run(run, p.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF, p.REDUCED_KERNEL_SIZE, cells, count)
        self._after_device_step()
        k = int(count.item())
        self._since_recompute += 1
        self._maybe_recompute()
        c = cells[:k].cpu().numpy().astype(np.int64)
        return np.asarray([c // self.nx, c % self.nx])

    def _maybe_recompute(self, st=None):
        if self.recompute_every and self._since_recompute >= self.recompute_every and self._incremental_exact():
            self._device_recompute(periodic=True, st=st)

    def _incremental_exact(self):
        """True when incremental updates reproduce a full recompute (no kernel truncation)."""
        if self.params.UPDATE_SCHEME == Scheme.NEEL:
            return True
        k = self.params.REDUCED_KERNEL_SIZE
        kern = self._dev.kern
        if k == 0 or kern is None:
            return True
        return min(k, self.nx - 1) >= kern["kwx"] and min(k, self.ny - 1) >= kern["kwy"]

    def calc_r(self, Q, as_scalar=True):
        """core.py:720-733 (cached per Q like the reference's lru_cache)."""
        key = (Q, as_scalar)
        if key not in self._calc_r_cache:
            with np.errstate(divide='ignore'):
                r = (8e-7 * self._momentSq / (Q * self.kBT)) ** (1 / 3)
            r = np.clip(r, 0, (self.x_max - self.x_min) + (self.y_max - self.y_min))
            self._calc_r_cache[key] = float(np.max(r)) if as_scalar else r
        return self._calc_r_cache[key]

    def _select_geometry(self, r):
        """Supergrid of _select_grid (core.py:505-540) for distance r: sizes and the occupied-cell
        pattern of one supercell for every unit-cell offset, uploaded once per (r, occupation)."""
        scheme = self.params.MULTISAMPLING_SCHEME
        if scheme not in ('grid', 'single', 'Poisson'):
            raise NotImplementedError(f"MULTISAMPLING_SCHEME='{scheme}' has no device implementation.")
        key = (float(r), scheme, self.params.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF,
               self.params.REDUCED_KERNEL_SIZE, bool(self.params.METROPOLIS_TIME))
        if self._sel_key == key:
            return self._sel
        sel = _lib.Select()
        sel.cutoff = self.params.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF
        sel.reduced_k = self.params.REDUCED_KERNEL_SIZE
        sel.metropolis_time = int(bool(self.params.METROPOLIS_TIME))
        t = self._dev.torch
        if scheme == 'single':
            sel.scheme = 1
            sel.nsc = 1
            self._sel_tensors = ()
        elif scheme == 'Poisson':
            # minimal distance r in cells along each axis (at least 1.5 cells, so snapped samples
            # stay distinct); background cells of r/sqrt(2) hold at most one sample each
            sel.scheme = 2
            sel.r_x = max(1.5, r / float(np.min(self.dx)))
            sel.r_y = max(1.5, r / float(np.min(self.dy)))
            sel.snx = int(self.nx / (sel.r_x * 0.70710678)) + 1  # one spare column/row: covers the grid
            sel.sny = int(self.ny / (sel.r_y * 0.70710678)) + 1  # whatever the FP32 cell size rounds to
            sel.nsc = sel.snx * sel.sny
            self._sel_tensors = ()
        else:
            sel.scheme = 0
            Rx = math.ceil(r / np.min(self.dx)) - 1
            Ry = math.ceil(r / np.min(self.dy)) - 1
            Rx = self.unitcell.x * math.ceil(Rx / self.unitcell.x)
            Ry = self.unitcell.y * math.ceil(Ry / self.unitcell.y)
            Rx, Ry = min(Rx, self.nx), min(Ry, self.ny)
            Rx, Ry = max(Rx, 2), max(Ry, 2)
            snx, sny = self.nx // Rx + 1, self.ny // Ry + 1
            ncols = len(range(0, min(snx, self.nx), 2))
            nrows = len(range(0, min(sny, self.ny), 2))
            offs, cells = [0], []
            for dy in range(self.unitcell.y):
                for dx in range(self.unitcell.x):
                    occ = self.occupation[dy:dy + Ry, dx:dx + Rx]
                    oy, ox = np.nonzero(occ)
                    if oy.size == 0:
                        raise ValueError("A supercell of the selection grid holds no magnet.")
                    cells.extend(((oy.astype(np.int64) << 16) | ox).tolist())
                    offs.append(len(cells))
            t_off = t.tensor(offs, dtype=t.int32, device=self._dev.device)
            t_cells = t.tensor(cells, dtype=t.int32, device=self._dev.device)
            self._sel_tensors = (t_off, t_cells)
            sel.pat_off, sel.pat_cells = t_off.data_ptr(), t_cells.data_ptr()
            sel.Rx, sel.Ry, sel.snx, sel.sny, sel.ncols, sel.nsc = Rx, Ry, snx, sny, ncols, ncols * nrows
        self._sel = sel
        self._sel_key = key
        return sel

    def progress(self, t_max: float = 1, MCsteps_max: float = 4., stepwise: bool = False, **kwargs):
        """core.py:736-751. Runs on the device until dt >= t_max or dMCsteps >= MCsteps_max."""
        if stepwise:
            return self._progress_stepwise(t_max=t_max, MCsteps_max=MCsteps_max, **kwargs)
        if self._T_any_zero:
            warnings.warn("T=0 somewhere in the domain, so no switch will be simulated to prevent DIV/0 errors.", stacklevel=2)
            # no update() advances t or MCsteps here; the reference would loop forever, so return at once
            return 0.0, 0.0
        self._pre_step()
        scheme = self.params.UPDATE_SCHEME
        if scheme == Scheme.WOLFF:
            return self._progress_wolff(t_max, MCsteps_max, **kwargs)
        attempt_freq = kwargs.get('attempt_freq', 1e10)
        # the run state is read once here and once after every launch: nothing else changes it in
        # between (the periodic recompute leaves t and the counters alone)
        st0 = self._state()
        t0, a0 = st0.t, st0.attempts
        begin = True
        st = st0
        if scheme == Scheme.NEEL:
            if kwargs.keys() - {'attempt_freq'}:
                raise TypeError(f"unexpected Neel keywords {sorted(kwargs.keys() - {'attempt_freq'})}")
            # A launch that runs all its steps (t_max infinite, the MC budget lasting past them: one
            # attempt per step) is queued without reading the state back; the host advances the
            # attempt count, and the state is read after a launch that may stop early, or at the end.
            t_cur, att = st.t, st.attempts
            unread = False
            while True:
                done_mc = (att - a0)
                if not (lower_than(t_cur - t0, t_max) and lower_than(att / self.n - a0 / self.n, MCsteps_max)):
                    break
                if math.isfinite(MCsteps_max):
                    chunk = max(1, int(math.ceil(MCsteps_max * self.n - done_mc)) + 1)
                else:
                    chunk = 4096
                if math.isfinite(t_max):
                    chunk = min(chunk, 4096)
                if self.recompute_every:
                    chunk = min(chunk, max(1, self.recompute_every - self._since_recompute))
                run = _lib.Run(t_max=float(t_max), mc_max=float(MCsteps_max), attempt_freq=float(attempt_freq),
                               max_steps=int(chunk), log=0, nblocks=0, begin=int(begin))
                team = self._neel_team_ok()
                if begin and not team:  # the team kernel takes the origin itself (hs_run_t.begin)
                    self._set_progress_origin(t0, a0, st0)
                self._dev.neel_run(run, persistent=team, nblocks=self._neel_nblocks())
                begin = False
                full = (math.isinf(t_max) and t_max > 0 and math.isfinite(MCsteps_max)
                        and chunk <= math.ceil(MCsteps_max * self.n - done_mc))
                if full:
                    att += chunk
                    unread = True
                    self._since_recompute += chunk
                    self._maybe_recompute()
                    continue
                st = self._state()
                unread = False
                if st.err:
                    raise _lib.HotspiceError(f"device Neel step failed (err={st.err})")
                t_cur, att = st.t, st.attempts
                self._since_recompute += int(st.steps_done)
                self._maybe_recompute(st)
            if unread:
                st = self._state()
                if st.err:
                    raise _lib.HotspiceError(f"device Neel step failed (err={st.err})")
        else:
            r = kwargs.get('r', 0)
            Q = kwargs.get('Q', 0.05)
            if kwargs.get('idx') is not None:  # every step samples the same list (core.py:757-758)
                sel = self._explicit_selection(kwargs['idx'])
                if sel is None:
                    raise ValueError("progress() with an empty idx never advances t or MCsteps")
                nblocks = self._metro_grid() if self.metro_blocks != 'auto' else 0
            else:
                r = self.calc_r(Q) if r == 0 else r
                nblocks = self._metro_nblocks(r)
                sel = self._select_geometry(r)
            while True:
                per_step = self._metro_rate or max(1.0, 0.25 * sel.nsc)
                if not (lower_than(st.t - t0, t_max) and lower_than(st.attempts / self.n - a0 / self.n, MCsteps_max)):
                    break
                if math.isfinite(MCsteps_max):
                    chunk = int(math.ceil((MCsteps_max * self.n - (st.attempts - a0)) / per_step)) + 1
                    chunk = max(1, min(chunk, 4096))
                else:
                    chunk = 1024
                if self.recompute_every and self._incremental_exact():
                    chunk = min(chunk, max(1, self.recompute_every - self._since_recompute))
                run = _lib.Run(t_max=float(t_max), mc_max=float(MCsteps_max), attempt_freq=float(attempt_freq),
                               max_steps=int(chunk), log=0, nblocks=nblocks, begin=int(begin))
                a_before = st.attempts  # (the launch takes the origin itself: hs_run_t.begin)
                self._dev.metropolis_run(sel, run)
                begin = False
                st = self._state()
                if st.err:
                    raise _lib.HotspiceError(f"device Metropolis step failed (err={st.err})")
                if st.steps_done > 0:  # samples per step, to size the next launch
                    self._metro_rate = max(1.0, (st.attempts - a_before) / st.steps_done)
                self._since_recompute += int(st.steps_done)
                self._maybe_recompute(st)
        self._after_device_step()
        return st.t - t0, st.attempts / self.n - a0 / self.n

    def _progress_wolff(self, t_max, MCsteps_max, **kwargs):
        """core.py:736-751 with _update_Wolff steps (which do not advance t), chunked on the device."""
        if kwargs:
            raise TypeError(f"unexpected Wolff keywords {sorted(kwargs)}")
        self._wolff_check()
        st0 = self._state()
        t0, a0 = st0.t, st0.attempts
        begin = True
        p = self.params
        while True:
            st = self._state()
            if not (lower_than(st.t - t0, t_max) and lower_than(st.attempts / self.n - a0 / self.n, MCsteps_max)):
                break
            chunk = 256
            if self.recompute_every and self._incremental_exact():
                chunk = min(chunk, max(1, self.recompute_every - self._since_recompute))
            run = _lib.Run(t_max=float(t_max), mc_max=float(MCsteps_max), attempt_freq=1e10, max_steps=int(chunk),
                           log=0, nblocks=0, begin=int(begin))
            if begin:
                self._set_progress_origin(t0, a0)
            self._dev.wolff_run(run, p.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF, p.REDUCED_KERNEL_SIZE)
            begin = False
            st = self._state()
            if st.err:
                raise _lib.HotspiceError(f"device Wolff step failed (err={st.err})")
            self._since_recompute += int(st.steps_done)
            self._maybe_recompute(st)
        self._after_device_step()
        st = self._state()
        return st.t - t0, st.attempts / self.n - a0 / self.n

    def _set_progress_origin(self, t0, a0, st=None):
        """Stores the origin (t0, attempts0) of a progress call in the run state (`st`: the current
        state, if read already; it is updated in place)."""
        st = self._state() if st is None else st
        st.t0, st.attempts0 = t0, a0
        self._dev.write_state(st)

    def _progress_stepwise(self, t_max: float = 1, MCsteps_max: float = 4., **kwargs):
        """core.py:753-760 (one device step per yield)."""
        t_0, MCsteps_0 = self.t, self.MCsteps
        while lower_than(dt := self.t - t_0, t_max) and lower_than(self.MCsteps - MCsteps_0, MCsteps_max):
            if self.params.UPDATE_SCHEME == Scheme.NEEL:
                self.update(t_max=(t_max - dt), **kwargs)
            else:
                self.update(**kwargs)
            yield
        return self.t - t_0, self.MCsteps - MCsteps_0

    # ---- zero-temperature relaxation ------------------------------------------------------------
    def minimize(self, ignore_barrier=True, verbose=False, _frac=0.3):
        """core.py:763-783 on the device: while some magnets have a barrier below -kT, switch a
        random ceil(_frac*n) of those n (numpy.random.choice, as the reference); otherwise switch
        the magnet with the smallest barrier while that is negative, each magnet at most 3 times.
        ignore_barrier=False ranks by the smaller E_barrier (the reference compares the pair of
        barriers there, which its indexing cannot use)."""
        self._pre_step()
        p = self.params
        visited = np.zeros_like(self.m)
        while True:
            n, amin, bmin = self._dev.minimize_scan(not ignore_barrier)
            if n != 0:  # magnets without any barrier: a random part of them
                cells = self._dev.min_cells[:n].cpu().numpy()
                rand = np.random.choice(n, size=math.ceil(n * _frac), replace=False)
                chosen = cells[rand]
            else:  # the one with the most energy to gain
                y, x = divmod(amin, self.nx)
                if bmin >= 0:
                    break
                visited[y, x] += 1
                if visited[y, x] >= 3:
                    break
                chosen = np.array([amin])
            if verbose:
                print(f"minimize: switching {chosen.size} magnet(s)")
            self._dev.flip_cells(chosen, p.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF, p.REDUCED_KERNEL_SIZE)
            self._after_device_step()
        st = self._state()
        if st.err:
            raise _lib.HotspiceError(f"device switch failed (err={st.err})")

    def _minimize_all(self, ignore_barrier=True, _simultaneous=False):
        """core.py:785-806 on the device (hs_relax_sweep): the subsets of cells two unit cells apart,
        in a random order (numpy.random.permutation, as the reference draws it), each switch every
        magnet of theirs that gains energy by switching (with ignore_barrier=False: that switches
        over its barrier, switch_energy/2 + E_B < 0), all at once, and update the fields."""
        if _simultaneous:
            warnings.warn("_simultaneous=True in _minimize_all(). I hope you know what you are doing, as this is very nonphysical ;)", stacklevel=2)
            step_x = step_y = 1
        else:
            step_x, step_y = self.unitcell.x * 2, self.unitcell.y * 2
        order = np.random.permutation(step_x * step_y)
        # a subset whose first cell lies outside the grid ends the sweep with the reference's
        # IndexError (core.py:797), after the subsets before it
        bad = [k for k, i in enumerate(order) if i // step_x >= self.ny or i % step_x >= self.nx]
        run = order[:bad[0]] if bad else order
        self._pre_step()
        p = self.params
        self._dev.relax_sweep(run, step_x, step_y, not ignore_barrier, p.SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF,
                              p.REDUCED_KERNEL_SIZE)
        self._after_device_step()
        if bad:
            i = int(order[bad[0]])
            raise IndexError(f"subset origin {divmod(i, step_x)} is outside the {self.ny}x{self.nx} grid")
        # the periodic FP64 recompute, unless truncated update_multiple windows are part of the result
        k, kern = p.REDUCED_KERNEL_SIZE, self._dev.kern
        if k == 0 or kern is None or (min(k, self.nx - 1) >= kern["kwx"] and min(k, self.ny - 1) >= kern["kwy"]):
            self._since_recompute += len(order)
            if self.recompute_every and self._since_recompute >= self.recompute_every:
                self._device_recompute()

    def relax(self, verbose=False, **kwargs):
        """core.py:808-845: _minimize_all until a sweep switches nothing, the state repeats with
        period 2, or sqrt(nx^2 + ny^2) sweeps (enough for a signal to cross the grid) are done."""
        n, n_max = 0, int(math.sqrt(self.nx ** 2 + self.ny ** 2))
        previous_states = [self.m.copy(), self._zeros, self._zeros]
        switches = self.switches
        while not np.allclose(previous_states[2], previous_states[0]) and n < n_max:
            n += 1
            self._minimize_all(**kwargs)
            if switches == self.switches:
                break
            switches = self.switches
            previous_states[2] = previous_states[1].copy()
            previous_states[1] = previous_states[0].copy()
            previous_states[0] = self.m.copy()
        if verbose:
            print(f"Used {n} (out of {n_max}) steps in relax().")

    # ---- observables (summed on the device: hs_binned_sums over one bin) ---------------------------
    def _m_sums(self):
        """(sum of m*ox, sum of m*oy) for in-plane lattices, (sum of m,) otherwise, over the grid."""
        self._sync_m()
        if getattr(self, "_obs_all", None) is None:
            w = np.stack([self.orientation[:, :, 0], self.orientation[:, :, 1]]) if self.in_plane else None
            self._obs_all = self._dev.bins([0, self.nx], [0, self.ny], w)
        return self._dev.binned_sums(self._obs_all).reshape(-1)

    @property
    def m_avg_x(self) -> float:
        """core.py:413-415: mean(m*ox) * nx*ny / n."""
        return float(self._m_sums()[0] / (self.nx * self.ny)) * self.nx * self.ny / self.n if self.in_plane else 0

    @property
    def m_avg_y(self) -> float:
        """core.py:416-418"""
        return float(self._m_sums()[1] / (self.nx * self.ny)) * self.nx * self.ny / self.n if self.in_plane else 0

    @property
    def m_avg(self) -> float:
        """core.py:419-421"""
        if not self.in_plane:
            return float(self._m_sums()[0] / (self.nx * self.ny))
        sx, sy = self._m_sums()[:2]
        mx = float(sx / (self.nx * self.ny)) * self.nx * self.ny / self.n
        my = float(sy / (self.nx * self.ny)) * self.nx * self.ny / self.n
        return (mx ** 2 + my ** 2) ** (1 / 2)

    @property
    def m_avg_angle(self) -> float:
        return math.atan2(self.m_avg_y, self.m_avg_x) if self.in_plane else 0. + (self.m_avg < 0) * math.pi

    @property
    def T_avg(self) -> float:
        return float(np.mean(self.T))

    @property
    def E_B_avg(self) -> float:
        return float(np.mean(self._E_B))

    @property
    def moment_avg(self) -> float:
        return float(np.mean(self.moment))

    def history_save(self, *, E_tot=None, t=None, T_avg=None, m_avg=None):
        self.history.entry(self)
        if E_tot is not None:
            self.history.E[-1] = float(E_tot)
        if t is not None:
            self.history.t[-1] = float(t)
        if T_avg is not None:
            self.history.T[-1] = float(T_avg)
        if m_avg is not None:
            self.history.m[-1] = float(m_avg)

    # ---- lattice hooks ---------------------------------------------------------------------------
    @abstractmethod
    def _get_angles(self):
        return np.zeros_like(self.ixx)

    @abstractmethod
    def _get_occupation(self):
        return np.ones_like(self.ixx)

    @abstractmethod
    def _get_groundstate(self):
        return 'random'

    @abstractmethod
    def _get_nearest_neighbors(self):
        return np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]])
