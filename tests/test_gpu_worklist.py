"""Metropolis steps on grids above the persistent-kernel size run as a selection launch and an
update launch over the worklist of tiles the step's switches reach. A tile missing from that list
would leave a switched cell unflipped or a field unchanged; the incremental FP32 fields would then
part from an FP64 recompute of the same spins by O(1). These tests step without intermediate
recomputes and require agreement to 1e-5 relative (float32 + Kahan drift is ~1e-7 here)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel_err(mm):
    E_inc = mm.E
    mm._device_recompute()
    E_ref = mm.E
    return float(np.max(np.abs(E_inc - E_ref)) / np.max(np.abs(E_ref)))


def _oop(n, energies, seed, T=600, PBC=False, **params):
    import hotspice_amd as hs
    rng = np.random.default_rng(seed)
    kT = hs.kB * T
    return hs.ASI.OOP_Square(1e-6, n, T=T, E_B=20 * kT * rng.normal(1, 0.05, (n, n)), moment=1.1e-15,
                             pattern=np.where(rng.random((n, n)) < .5, -1., 1.), seed=seed, energies=energies,
                             params=hs.SimParams(UPDATE_SCHEME='Metropolis', **params), PBC=PBC,
                             recompute_every=None)


def test_full_kernel_lists_every_tile():
    """264 x 264 cells with the full 527 x 527 kernel: every switch reaches all 297 tiles of
    32 x 8 cells, more than a switch marks one by one, so the selection lists the whole grid."""
    import hotspice_amd as hs
    mm = _oop(264, [hs.DipolarEnergy()], 41, REDUCED_KERNEL_SIZE=0)
    s0 = mm.switches
    mm.progress(t_max=math.inf, MCsteps_max=0.02, r=40e-6)
    assert mm.switches > s0
    assert _rel_err(mm) < 1e-5


@pytest.mark.parametrize("pbc", [False, True])
def test_truncated_windows_and_exchange(pbc):
    """160 x 160 cells, truncated 17 x 17 windows (wrapped with PBC) plus the exchange stencil."""
    import hotspice_amd as hs
    kT = hs.kB * 600
    mm = _oop(160, [hs.DipolarEnergy(cutoff_range=8e-6), hs.ExchangeEnergy(J=0.2 * kT)], 42 + pbc, PBC=pbc,
              REDUCED_KERNEL_SIZE=8, SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0)
    s0 = mm.switches
    mm.progress(t_max=math.inf, MCsteps_max=1.0, r=12e-6)
    assert mm.switches > s0
    assert _rel_err(mm) < 1e-5


def test_exchange_only():
    import hotspice_amd as hs
    kT = hs.kB * 600
    mm = _oop(120, [hs.ExchangeEnergy(J=0.5 * kT)], 44)
    s0 = mm.switches
    mm.progress(t_max=math.inf, MCsteps_max=1.0, r=4e-6)
    assert mm.switches > s0
    assert _rel_err(mm) < 1e-5
