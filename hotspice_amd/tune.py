"""Autotuner for the launch shapes of the stepping kernels.

The Neel step runs as a team of G co-resident 1024-thread workgroups per instance; the best G
trades the per-workgroup work of a step (cells per CU) against the cost of exchanging the
winner between workgroups, and depends on the grid size, the sublattices and the device.
`tune_neel_blocks(mm)` times short persistent runs for a few team sizes on the instance itself,
keeps the fastest and caches it per geometry, so later instances of the same shape start with
it. The instance's device state (spins, fields, counters, Philox step) is saved before the trial
runs and restored afterwards, so tuning does not change the simulation, and every team size
gives bit-identical results anyway.

`neel_blocks='auto'` on a Magnets instance runs this before its first Neel launch.
"""
import math
import time

_CACHE = {}


def _key(mm):
    return (mm.nx, mm.ny, mm.unitcell.x, mm.unitcell.y, bool(mm.USE_PERP_ENERGY), mm._dev.kern["kwx"] if mm._dev.kern
            else 0, mm._dev.kern["kwy"] if mm._dev.kern else 0)


def neel_candidates(mm, max_workgroups=256):
    """Team sizes worth trying: around one magnet per owner thread (n/512 workgroups) down to 16
    magnets per owner thread, at most the co-resident limit."""
    base = max(1, math.ceil(mm.n / 512))
    out = []
    for f in (2, 1, 0.75, 0.5, 0.25):
        g = max(1, min(max_workgroups, int(round(base * f))))
        if math.ceil(mm.n / (g * 512)) <= 16 and g not in out:
            out.append(g)
    return out or [max(1, min(max_workgroups, base))]


def _snapshot(dev):
    names = ("m", "E", "Ec", "Ep", "Epc", "state")
    return {n: getattr(dev, n).clone() for n in names if getattr(dev, n, None) is not None}


def _restore(dev, snap):
    for n, t in snap.items():
        getattr(dev, n).copy_(t)


def tune_neel_blocks(mm, candidates=None, steps=256, reuse=True):
    """Times `steps` Neel steps for each candidate team size on `mm` (state restored
    afterwards), sets mm.neel_blocks to the fastest and returns {G: microseconds per step}."""
    import torch
    key = _key(mm)
    if reuse and key in _CACHE:
        mm.neel_blocks = _CACHE[key]
        return {}
    props = torch.cuda.get_device_properties(mm._dev.device)
    cands = candidates or neel_candidates(mm, props.multi_processor_count)
    snap = _snapshot(mm._dev)
    since, every = mm._since_recompute, mm.recompute_every
    mm.recompute_every = None
    timing = {}
    try:
        for g in cands:
            mm.neel_blocks = g
            mm.progress(t_max=math.inf, MCsteps_max=8 / mm.n)  # warm-up (code object, caches)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            mm.progress(t_max=math.inf, MCsteps_max=steps / mm.n)
            torch.cuda.synchronize()
            timing[g] = (time.perf_counter() - t0) / steps * 1e6
            _restore(mm._dev, snap)
    finally:
        _restore(mm._dev, snap)
        mm._after_device_step()
        mm._since_recompute, mm.recompute_every = since, every
    best = min(timing, key=timing.get)
    mm.neel_blocks = best
    _CACHE[key] = best
    return timing
