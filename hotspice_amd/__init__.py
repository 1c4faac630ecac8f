"""hotspice_amd: the Monte Carlo stepping path of Hotspice on AMD MI355X (gfx950).

Drop-in for the stepping API of bvwaeyen/Hotspice: ``Magnets.update()`` / ``progress()``
under ``SimParams.UPDATE_SCHEME`` in {Neel, Metropolis}, with ``DipolarEnergy`` (or
``energies.DiMonopolarEnergy``), ``ZeemanEnergy`` and ``ExchangeEnergy``. All stepping runs in hand-written HIP kernels
(libhotspice_hip.so, C-ABI in include/hotspice_hip.h); PyTorch provides device memory.

    import hotspice_amd as hotspice
    mm = hotspice.ASI.IP_Pinwheel(1e-6, 256, T=300, E_B=..., moment=2.37e-16,
                                  energies=[hotspice.DipolarEnergy(), hotspice.ZeemanEnergy(1e-3, 0.39)])
    mm.progress(t_max=float('inf'), MCsteps_max=10)
"""
from . import utils
from .core import History, Magnets, Scheme, SimParams, kB
from .energies import DipolarEnergy, Energy, ExchangeEnergy, ZeemanEnergy
from . import energies
from . import ASI
from . import io

__all__ = ['ASI', 'energies', 'io', 'utils', 'kB', 'SimParams', 'Magnets', 'Scheme', 'History', 'Energy',
           'ZeemanEnergy', 'DipolarEnergy', 'ExchangeEnergy']
