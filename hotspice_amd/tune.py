"""Autotuner for the launch shapes of the stepping kernels, with a cache on disk.

Two shapes depend on the instance, the workload and the device, and are picked by timing them on
the instance itself:
  * the Neel team size G (`Magnets.neel_blocks`): a team of G co-resident 1024-thread
    workgroups trades the per-workgroup work of a step (magnets per owner thread) against the
    cost of exchanging the winner between workgroups;
  * the Metropolis update grid (`Magnets.metro_blocks`, hs_run_t.nblocks): the workgroups that
    work through a step's tile worklist; more workgroups take fewer tiles each in turn, fewer
    keep more of them resident.
Each tuner saves the instance's device state (spins, fields, compensations, counters, Philox
step) before its trial runs and restores it afterwards, so tuning does not change the simulation;
every candidate gives bit-identical results anyway. The choice is cached per geometry and device,
in memory and in a JSON file (HOTSPICE_AMD_TUNE_CACHE, default ~/.cache/hotspice_amd/tune.json),
so later instances and later processes of the same shape start with it.

`neel_blocks='auto'` on a Magnets instance runs the Neel tuner before its first Neel launch;
`metro_blocks='auto'` runs the Metropolis tuner before its first launch for each selection
geometry (sampling distance, scheme).
"""
import json
import math
import os
import time

_CACHE = {}
_LOADED = False


def cache_path():
    return os.environ.get("HOTSPICE_AMD_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "hotspice_amd", "tune.json"))


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    try:
        with open(cache_path()) as f:
            _CACHE.update(json.load(f))
    except (OSError, ValueError):
        pass


def _store(key, value):
    _CACHE[key] = value
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = f"{path}.{os.getpid()}.tmp"
        with open(tmp, "w") as f:
            json.dump(_CACHE, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home: the choice stays in memory


def _device_name(mm):
    import torch
    return torch.cuda.get_device_properties(mm._dev.device).name


def _key(kind, mm, *extra):
    kern = mm._dev.kern
    parts = (kind, mm.nx, mm.ny, mm.unitcell.x, mm.unitcell.y, bool(mm.USE_PERP_ENERGY), bool(mm.PBC),
             kern["kwx"] if kern else 0, kern["kwy"] if kern else 0, *extra, _device_name(mm))
    return "|".join(str(p) for p in parts)


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


def metro_candidates(mm):
    """Update grids worth trying: 1024 to 16384 workgroups, at most one per tile (32 x 8 cells)."""
    tiles = math.ceil(mm.nx / 32) * math.ceil(mm.ny / 8)
    out = sorted({min(tiles, g) for g in (1024, 2048, 4096, 8192, 16384)})
    return out


def _snapshot(dev):
    names = ("m", "E", "Ec", "Ep", "Epc", "E64", "state")
    return {n: getattr(dev, n).clone() for n in names if getattr(dev, n, None) is not None}


def _restore(dev, snap):
    for n, t in snap.items():
        getattr(dev, n).copy_(t)


def _timed(mm, candidates, run_one):
    """{candidate: seconds} of run_one(candidate), each from the same saved state."""
    import torch
    snap = _snapshot(mm._dev)
    since, every = mm._since_recompute, mm.recompute_every
    mm.recompute_every = None
    timing = {}
    try:
        for c in candidates:
            run_one(c)  # warm-up (code object, caches)
            torch.cuda.synchronize()
            _restore(mm._dev, snap)
            t0 = time.perf_counter()
            run_one(c)
            torch.cuda.synchronize()
            timing[c] = time.perf_counter() - t0
            _restore(mm._dev, snap)
    finally:
        _restore(mm._dev, snap)
        mm._after_device_step()
        mm._since_recompute, mm.recompute_every = since, every
    return timing


def tune_neel_blocks(mm, candidates=None, steps=256, reuse=True):
    """Times `steps` Neel steps for each candidate team size on `mm` (state restored afterwards),
    sets mm.neel_blocks to the fastest and returns {G: microseconds per step} ({} when the cached
    choice was taken)."""
    import torch
    _load()
    key = _key("neel", mm)
    if reuse and key in _CACHE:
        mm.neel_blocks = int(_CACHE[key])
        return {}
    props = torch.cuda.get_device_properties(mm._dev.device)
    cands = candidates or neel_candidates(mm, props.multi_processor_count)

    def run_one(g):
        mm.neel_blocks = g
        mm.progress(t_max=math.inf, MCsteps_max=steps / mm.n)

    timing = {g: s / steps * 1e6 for g, s in _timed(mm, cands, run_one).items()}
    best = min(timing, key=timing.get)
    mm.neel_blocks = best
    _store(key, best)
    return timing


def tune_metro_blocks(mm, r, candidates=None, steps=64, reuse=True):
    """Times `steps` Metropolis steps at sampling distance r for each candidate update grid on
    `mm` (state restored afterwards), records the fastest as the instance's grid for this
    selection geometry (used when mm.metro_blocks == 'auto') and returns {workgroups:
    microseconds per step} ({} when the cached choice was taken)."""
    from . import _lib
    _load()
    sel = mm._select_geometry(r)
    key = _key("metropolis", mm, sel.scheme, sel.nsc, sel.Rx, sel.Ry, sel.reduced_k)
    if reuse and key in _CACHE:
        mm._metro_tuned[mm._sel_key] = int(_CACHE[key])
        return {}
    cands = candidates or metro_candidates(mm)

    def run_one(g):
        mm._pre_step()
        run = _lib.Run(t_max=math.inf, mc_max=math.inf, attempt_freq=1e10, max_steps=steps, log=0, nblocks=int(g),
                       begin=1, first_free=0)
        mm._dev.metropolis_run(sel, run)

    timing = {g: s / steps * 1e6 for g, s in _timed(mm, cands, run_one).items()}
    best = min(timing, key=timing.get)
    mm._metro_tuned[mm._sel_key] = best
    _store(key, best)
    return timing
