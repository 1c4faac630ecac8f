"""Energy plugins of the stepping path: ZeemanEnergy, DipolarEnergy, DiMonopolarEnergy, ExchangeEnergy.

Same classes and members as hotspice/energies.py (Energy contract 22-83, ZeemanEnergy 86-140,
DipolarEnergy 143-391, DiMonopolarEnergy 394-751, ExchangeEnergy 754-791). Unlike the reference, the per-magnet fields
live on the GPU as ONE combined FP32 local field with Kahan compensation (the sum of every
energy's E), updated by the HIP kernels of libhotspice_hip.so. The per-energy arrays E and
E_perp are reconstructed on demand (FP64) from that combined field and the closed forms of
the Zeeman and exchange terms, so user code that reads ``energy.E`` keeps working.
"""
from __future__ import annotations

import math
from abc import ABC, abstractmethod

import numpy as np

from .utils import as_2D_array


class Energy(ABC):
    """hotspice/energies.py:22-83. Subclasses describe their terms to the device instance."""

    def __init__(self):
        pass

    def energy_switch(self, indices2D=None):
        """energies.py:29-35"""
        E = self.E
        return -2 * E if indices2D is None else -2 * E[indices2D[0], indices2D[1]]

    def initialize(self, mm):
        """energies.py:37-42: attach to `mm`; the device fields are recomputed by Magnets."""
        self.mm = mm
        self._initialize()

    @abstractmethod
    def _initialize(self):
        ...

    def update(self):
        """Full FP64 recompute of all device fields (energies.py:51-55 contract)."""
        self.mm._device_recompute()

    def update_single(self, index2D):
        self.mm._device_recompute()

    def update_multiple(self, indices2D):
        self.mm._device_recompute()

    @property
    def E(self):
        return self.mm._energy_component(self, perp=False)

    @property
    def E_perp(self):
        return self.mm._energy_component(self, perp=True)

    @property
    @abstractmethod
    def E_tot(self):
        ...

    @property
    def shortname(self):
        return type(self).__name__.lower().replace("energy", "")


class ZeemanEnergy(Energy):
    """energies.py:86-140. A uniform or 2D external field; set_field is the served 'field step'."""

    def __init__(self, magnitude=0, angle=0):
        self._magnitude = magnitude
        self._angle = angle

    def _initialize(self):
        pass

    def field_components(self):
        """(Bx, By) as scalars (uniform field) or (ny, nx) arrays; OOP: (B, 0)."""
        mm = self.mm
        mag, ang = self._magnitude, self._angle
        uniform = np.asarray(mag).size == 1 and np.asarray(ang).size == 1
        if uniform:
            mag, ang = float(np.asarray(mag).reshape(-1)[0]), float(np.asarray(ang).reshape(-1)[0])
            if mm.in_plane:
                return mag * math.cos(ang), mag * math.sin(ang), True
            return mag, 0.0, True
        magA, angA = as_2D_array(mag, mm.shape), as_2D_array(ang, mm.shape)
        if mm.in_plane:
            return magA * np.cos(angA), magA * np.sin(angA), False
        return magA.copy(), np.zeros(mm.shape), False

    def set_field(self, magnitude=None, angle=None):
        """energies.py:101-114: the combined device field swaps its Zeeman term in place."""
        old = None
        if hasattr(self, "mm"):
            old = self.field_components()
        if magnitude is not None:
            self._magnitude = magnitude
        if angle is not None:
            self._angle = angle
        if old is not None:
            self.mm._device_zeeman_changed(self, old)

    @property
    def magnitude(self):
        return as_2D_array(self._magnitude, self.mm.shape)

    @magnitude.setter
    def magnitude(self, value):
        self.set_field(magnitude=value)

    @property
    def angle(self):
        return as_2D_array(self._angle, self.mm.shape)

    @angle.setter
    def angle(self, value):
        self.set_field(angle=value)

    @property
    def E_factor(self):
        mm = self.mm
        bx, by, _ = self.field_components()
        if mm.in_plane:
            return -mm.moment * (bx * mm.orientation[:, :, 0] + by * mm.orientation[:, :, 1])
        return -mm.moment * bx

    @property
    def E_tot(self):
        return float(np.sum(self.E))


class DipolarEnergy(Energy):
    """energies.py:143-391. Kernels are built on the device (hs_build_dipolar_kernels)."""

    def __init__(self, prefactor: float = 1, decay_exponent: float = -3, cutoff_range: float = math.inf):
        self._prefactor = prefactor
        self.decay_exponent = decay_exponent
        self._cutoff_range = cutoff_range

    def _initialize(self):
        self.mm._device_build_kernels(self)

    @property
    def prefactor(self):
        return self._prefactor

    @prefactor.setter
    def prefactor(self, value):
        """energies.py:160-164 (E scales with the prefactor: recomputed exactly in FP64)."""
        self._prefactor = value
        if hasattr(self, "mm"):
            self.mm._device_params_changed()
            self.mm._device_recompute()

    @property
    def cutoff_range(self):
        return self._cutoff_range

    @cutoff_range.setter
    def cutoff_range(self, value):
        """energies.py:169-174"""
        self._cutoff_range = value
        if hasattr(self, "mm"):
            self._initialize()
            self.mm._device_params_changed()
            self.mm._device_recompute()

    @property
    def unitcell(self):
        return self.mm.unitcell

    @property
    def kernel_unitcell_indices(self):
        """(uy, ux) unit-cell site -> kernel index, -1 for empty sites (energies.py:198)."""
        mm = self.mm
        idx = -np.ones((mm.unitcell.y, mm.unitcell.x), dtype=int)
        for n, (y, x) in enumerate(mm._kernel_sites):
            idx[y, x] = n
        return idx

    @property
    def kernel_unitcell(self):
        """(n_sub, 2ny-1, 2nx-1) FP64 kernels (the stored window, zero-padded to full size)."""
        return self.mm._dev.kernels_full("K", self.mm.nx, self.mm.ny)

    @property
    def kernel_perpself_unitcell(self):
        return self.mm._dev.kernels_full("Kps", self.mm.nx, self.mm.ny)

    @property
    def kernel_perpother_unitcell(self):
        return self.mm._dev.kernels_full("Kpo", self.mm.nx, self.mm.ny)

    @property
    def E_tot(self):
        return float(np.sum(self.E) / 2)

    def get_NN_interaction(self):
        """energies.py:372-387"""
        largest = 0
        NN = np.asarray(self.mm._get_nearest_neighbors())
        for kernel in self.kernel_unitcell:
            ny, nx = kernel.shape
            my, mx = ny // 2, nx // 2
            dy, dx = NN.shape[0] // 2, NN.shape[1] // 2
            value = np.max(np.abs(kernel[my - dy:my + dy + 1, mx - dx:mx + dx + 1] * NN))
            largest = max(largest, value)
        return largest * self.prefactor * (self.mm.moment_avg ** 2)

    def set_NN_interaction(self, value):
        """energies.py:389-391"""
        self.prefactor = value / self.get_NN_interaction()


class DiMonopolarEnergy(DipolarEnergy):
    """energies.py:394-751. Every magnet is a dumbbell of two opposite magnetic charges
    q = moment/d, d apart along the magnet's angle; the interaction is the Coulomb-like sum of
    the four charge pairs. The kernels are built on the device (hs_build_dimonopolar_kernels)
    and the stepping kernels are the DipolarEnergy ones, with the pair term
    prefactor * K * m_source * m_receiver * q_receiver^2 (energies.py:654-659, 669-671, 707-710).
    The charges are taken from the moment when the energy is initialised, as in the reference.
    Replaces a DipolarEnergy of the same Magnets (the device holds one pair-interaction kernel set).
    """

    def __init__(self, prefactor=1, d=200e-9, small_d=200e-9):
        self._prefactor = prefactor
        self.d = d
        self.small_d = small_d

    def _initialize(self):
        """energies.py:406-415; the kernel construction (416-641) runs on the device."""
        mm = self.mm
        if not mm.in_plane:
            raise ValueError("DiMonopolarEnergy needs an in-plane ASI (magnet angles).")
        ux, uy = mm.unitcell.x, mm.unitcell.y
        if isinstance(self.d, np.ndarray):
            self.dist_unitcell = np.asarray(self.d, dtype=float)
            if self.dist_unitcell.shape != (uy, ux):
                raise ValueError(f"d must be a scalar or a ({uy}, {ux}) unit-cell array.")
        else:
            self.dist_unitcell = as_2D_array(self.d, (uy, ux))
        self.dist = np.tile(self.dist_unitcell, (math.ceil(mm.ny / uy), math.ceil(mm.nx / ux)))
        self.dist_exact = self.dist[:mm.ny, :mm.nx]
        self.charge = mm.moment / self.dist_exact  # [Am]
        self.charge_sq = self.charge ** 2
        mm._device_build_kernels(self)

    @property
    def cutoff_range(self):
        return math.inf


class ExchangeEnergy(Energy):
    """energies.py:754-791. Folded into the same incremental device update as a local stencil."""

    def __init__(self, J=1):
        self._J = J

    @property
    def J(self):
        return self._J

    @J.setter
    def J(self, value):
        self._J = value
        if hasattr(self, "mm"):
            self.mm._device_params_changed()
            self.mm._device_recompute()

    def _initialize(self):
        self.local_interaction = np.asarray(self.mm._get_nearest_neighbors())

    @property
    def E_tot(self):
        return float(np.sum(self.E) / 2)
