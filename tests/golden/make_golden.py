"""Writes replayed-trajectory fixtures from the reference implementation.

Runs ONLY in the build container, where the reference tree exists at /root/reference; the
fixtures it writes (tests/golden/*.npz) are data: inputs and the reference's outputs.

What is pinned here and why an import is needed: the exact trajectory semantics of
Magnets._update_Neel (hotspice/core.py:679-709), Magnets._update_Metropolis with grid
selection (core.py:456-558, 659-677), the 'single' scheme and the grid's fallback to it
(_select_single, core.py:489-494, 555-558), Magnets.progress with its two stop predicates and
the t_max - dt hand-off of Neel steps (core.py:736-760), and the update_single /
update_multiple branch rule (hotspice/energies.py:310-366). Closed forms and invariants pin the
kernels, barriers and statistics, but not these step-by-step trajectories, so they come from
the reference itself.

How: the reference's random draws are replaced by the keyed Philox streams of oracle/philox.py
(the same streams the HIP kernels draw), so the reference's trajectory is the one the build
must reproduce:
  mm.rng.exponential(scale, size)   -> scale * -ln(u_cell),  u from TAG_NEEL   (core.py:692, 698)
  mm.rng.random(size)               -> u per sample cell,    TAG_MTIME/ACCEPT  (core.py:668, 671)
  numpy.random.randint(...)         -> TAG_SEL slots 0/1 (offsets), 2 (single pick) (core.py:493, 527-529)
  numpy.random.choice(cnt, n)       -> floor(u_j * cnt) per supercell, TAG_CHOICE  (core.py:538)
Only the off-path modules colorama and IPython (imported by hotspice/utils.py) are stubbed;
hotspice/__init__.py is bypassed so the GUI / experiments modules are not imported.

Usage:  python tests/golden/make_golden.py [case ...]   (default: every case)
"""
import math
import os
import sys
import types

import numpy as np

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import philox  # noqa: E402


def import_reference():
    os.environ["HOTSPICE_USE_GPU"] = "False"
    for name in ("colorama",):
        m = types.ModuleType(name)
        m.init = lambda *a, **k: None
        m.Fore = m.Style = m.Back = types.SimpleNamespace(__getattr__=lambda *_: "")
        sys.modules[name] = m
    ip = types.ModuleType("IPython")
    ipt = types.ModuleType("IPython.terminal")
    ipe = types.ModuleType("IPython.terminal.embed")
    ipe.InteractiveShellEmbed = object
    sys.modules.update({"IPython": ip, "IPython.terminal": ipt, "IPython.terminal.embed": ipe})
    pkg = types.ModuleType("hotspice")
    pkg.__path__ = [os.path.join(REF, "hotspice")]
    sys.modules["hotspice"] = pkg
    import importlib
    core = importlib.import_module("hotspice.core")
    energies = importlib.import_module("hotspice.energies")
    ASI = importlib.import_module("hotspice.ASI")
    return core, energies, ASI


class Replay:
    """Serves Philox uniforms to the reference in its own consumption order."""

    def __init__(self, mm, seed):
        self.mm, self.seed = mm, seed
        self.step = 0
        self.attempt = -1
        self.call = 0
        self.in_single = False
        self.last_pos = None
        self.random_calls = 0
        self.samples = []
        self.singles = 0  # _select_single calls (the 'single' scheme, or the grid's fallback)
        outer = self

        class RNG:
            def exponential(self, scale=1.0, size=None):
                u = philox.neel_uniforms(outer.seed, outer.step, outer.mm.nx * outer.mm.ny).reshape(size)
                return scale * (-np.log(u))

            def random(self, size=None):
                pos = outer.last_pos
                cells = pos[0] * outer.mm.nx + pos[1]
                mtime = outer.mm.params.METROPOLIS_TIME and outer.random_calls == 0
                outer.random_calls += 1
                if mtime:
                    return philox.mtime_uniforms(outer.seed, outer.step, cells)
                return philox.accept_uniforms(outer.seed, outer.step, cells)

        mm.rng = RNG()
        orig_select = mm.select
        orig_single = mm._select_single

        def select(**kw):
            outer.attempt += 1
            outer.call = 0
            pos = orig_select(**kw)
            outer.last_pos = pos
            return pos

        def single(**kw):
            outer.in_single = True
            outer.singles += 1
            try:
                return orig_single(**kw)
            finally:
                outer.in_single = False

        mm.select = select
        mm._select_single = single

    def randint(self, *args):
        u = philox.sel_uniforms(self.seed, self.step, self.attempt)
        if self.in_single:
            v = u[2]
        else:
            v = u[self.call]
            self.call += 1
        return philox.randint(v, *args)

    def choice(self, a, size=None):
        u = philox.choice_uniforms(self.seed, self.step, self.attempt, size)
        return np.floor(u * a).astype(np.int64)

    def begin_step(self):
        self.attempt = -1
        self.random_calls = 0
        self.last_pos = None

    def end_step(self):
        self.step += 1


def run_case(core, energies, ASI, spec):
    """Builds the reference Magnets for `spec`, replays `nsteps` updates, returns arrays."""
    rng = np.random.default_rng(spec["seed"])
    cls = getattr(ASI, spec["lattice"])
    kw = dict(spec.get("kw", {}))
    # geometry first (to size the pattern and E_B arrays), with a throwaway uniform pattern
    probe = cls(spec["a"], spec["n"], pattern="uniform", energies=[], **{k: v for k, v in kw.items() if k in ("nx", "ny", "PBC")})
    shape = probe.m.shape
    occ = probe.occupation
    pattern = np.where(rng.random(shape) < 0.5, -1.0, 1.0)
    kT = core.kB * spec["T"]
    if spec["EB_dist"] == "normal":
        E_B = spec["EB_kT"] * kT * rng.normal(1, 0.05, size=shape)
    else:
        E_B = spec["EB_kT"] * kT * np.exp(0.05 * rng.normal(0, 1, size=shape))
    ens = []
    for e in spec["energies"]:
        if e[0] == "dipolar":
            ens.append(energies.DipolarEnergy(**e[1]))
        elif e[0] == "zeeman":
            ens.append(energies.ZeemanEnergy(**e[1]))
        elif e[0] == "exchange":
            ens.append(energies.ExchangeEnergy(J=e[1]["J_kT"] * kT))
    params = core.SimParams(**spec["params"])
    # a temperature that varies over the grid: T * (1 + T_ramp * (x / (nx-1) - 1/2))
    T = spec["T"]
    if "T_ramp" in spec:
        T = spec["T"] * (1 + spec["T_ramp"] * np.linspace(-0.5, 0.5, shape[1])[None, :] * np.ones((shape[0], 1)))
    # An array pattern cannot be passed to the constructor (core.py:172 reads self.shape before
    # self.m exists), so the state is set afterwards with initialize_m (core.py:163-180).
    mm = cls(spec["a"], spec["n"], T=T, E_B=E_B, moment=spec["moment"], pattern="uniform",
             energies=ens, params=params, **kw)
    mm.initialize_m(pattern)
    rep = Replay(mm, spec["seed"])
    saved = (np.random.randint, np.random.choice)
    np.random.randint, np.random.choice = rep.randint, rep.choice
    # every update() -- called directly, or by progress() through self.update (core.py:757-758)
    # -- is one replay step; its switches are logged with the step index
    log = dict(y=[], x=[], step=[])
    orig_update = mm.update

    def update(*args, **kwargs):
        rep.begin_step()
        r = orig_update(*args, **kwargs)
        if r is not None:
            arr = np.asarray(r).reshape(2, -1)
            log["y"].extend(arr[0].tolist())
            log["x"].extend(arr[1].tolist())
            log["step"].extend([rep.step] * arr.shape[1])
        rep.end_step()
        return r

    mm.update = update
    out = dict(pattern=pattern, E_B=E_B, occupation=occ, m0=mm.m.copy(), E0=mm.E.copy(),
               **({"T": np.asarray(T, dtype=float)} if "T_ramp" in spec else {}),
               Eperp0=(mm.E_perp.copy() if mm.USE_PERP_ENERGY else np.zeros(shape)),
               E_tot0=np.float64(mm.E_tot))
    E_dip0 = mm.get_energy("dipolar").E.copy() if mm.has_energy("dipolar") else np.zeros(shape)
    out["Edip0"] = E_dip0
    calls = spec.get("calls")
    nsteps = len(calls) if calls else spec["nsteps"]
    t_hist = np.zeros(nsteps)
    att_hist = np.zeros(nsteps, dtype=np.int64)
    snaps = []
    ret = np.zeros((nsteps, 2))
    steps_hist = np.zeros(nsteps, dtype=np.int64)
    sw_hist = np.zeros(nsteps, dtype=np.int64)
    m_hist = []
    try:
        for s in range(nsteps):
            if calls:  # one progress() call per entry (core.py:736-760)
                ret[s] = mm.progress(**calls[s])
                m_hist.append(mm.m.copy())
            else:
                mm.update(**spec.get("update_kw", {}))
            t_hist[s] = mm.t
            att_hist[s] = mm.attempted_switches
            steps_hist[s] = rep.step
            sw_hist[s] = mm.switches
            if (s + 1) % spec["snap_every"] == 0:
                snaps.append((s, mm.E.copy(), mm.E_perp.copy() if mm.USE_PERP_ENERGY else np.zeros(shape), mm.m.copy()))
    finally:
        np.random.randint, np.random.choice = saved
    out.update(sw_y=np.array(log["y"], dtype=np.int64), sw_x=np.array(log["x"], dtype=np.int64),
               sw_step=np.array(log["step"], dtype=np.int64), t=t_hist, attempts=att_hist,
               snap_steps=np.array([s[0] for s in snaps]), snap_E=np.array([s[1] for s in snaps]),
               snap_Eperp=np.array([s[2] for s in snaps]), snap_m=np.array([s[3] for s in snaps]),
               m_final=mm.m.copy(), E_final=mm.E.copy(), E_tot_final=np.float64(mm.E_tot),
               switches=np.int64(mm.switches))
    if spec.get("count_singles"):
        out["singles"] = np.int64(rep.singles)
    if calls:  # per progress() call: its return value, the update() steps and switches so far, m
        out.update(ret=ret, steps=steps_hist, switches_hist=sw_hist, m_hist=np.array(m_hist))
    return out


CASES = {
    # Neel, in-plane pinwheel, perp channels on (m_perp_factor default 1), Zeeman + dipolar
    "neel_pinwheel": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, moment=2.37e-16, EB_kT=60, EB_dist="lognormal",
                          energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=math.pi / 8))],
                          params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=np.inf), nsteps=200, snap_every=50, seed=11),
    # Neel, OOP square, parabolic barrier, dipolar + Zeeman + exchange
    "neel_oop_parabolic": dict(lattice="OOP_Square", a=1e-6, n=20, T=300, moment=1.1e-15, EB_kT=20, EB_dist="normal",
                               energies=[("dipolar", {}), ("zeeman", dict(magnitude=2e-5)), ("exchange", dict(J_kT=0.5))],
                               params=dict(UPDATE_SCHEME="Néel", ENERGY_BARRIER_METHOD="parabolic"),
                               update_kw=dict(t_max=np.inf), nsteps=200, snap_every=50, seed=12),
    # Neel, Kagome with exchange (XY projection) and a finite t_max (no-switch steps, core.py:701-704)
    "neel_kagome_tmax": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16, EB_kT=40, EB_dist="normal",
                             energies=[("dipolar", {}), ("zeeman", dict(magnitude=5e-4, angle=0.3)), ("exchange", dict(J_kT=0.3))],
                             params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=1e-3), nsteps=120, snap_every=40, seed=13),
    # Neel with a temperature that varies over the grid: the grid's minimum barrier reorders the
    # waiting times (core.py:686-699); pinwheel with perp channels, OOP square with parabolic
    # barriers and a finite t_max
    "neel_pinwheel_tramp": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, T_ramp=0.6, moment=2.37e-16, EB_kT=60,
                                EB_dist="lognormal", energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.4))],
                                params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=np.inf), nsteps=200, snap_every=50,
                                seed=14),
    "neel_oop_tramp_parabolic": dict(lattice="OOP_Square", a=1e-6, n=20, T=300, T_ramp=0.8, moment=1.1e-15, EB_kT=20,
                                     EB_dist="normal", energies=[("dipolar", {}), ("exchange", dict(J_kT=0.3))],
                                     params=dict(UPDATE_SCHEME="Néel", ENERGY_BARRIER_METHOD="parabolic"),
                                     update_kw=dict(t_max=1e-3), nsteps=200, snap_every=50, seed=15),
    # Metropolis, OOP square, full kernel, grid selection r=4a, sum branch (<= 20 per sublattice)
    "metro_oop_sum": dict(lattice="OOP_Square", a=1e-6, n=32, T=300, moment=1.1e-15, EB_kT=20, EB_dist="normal",
                          energies=[("dipolar", {})], params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0),
                          update_kw=dict(r=4e-6), nsteps=150, snap_every=50, seed=21),
    # Metropolis, OOP square, convolution branch with a truncated 11x11 window (cutoff 2)
    "metro_oop_conv": dict(lattice="OOP_Square", a=1e-6, n=32, T=600, moment=1.1e-15, EB_kT=20, EB_dist="normal",
                           energies=[("dipolar", {}), ("exchange", dict(J_kT=0.2))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=5,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=2),
                           update_kw=dict(r=3e-6), nsteps=150, snap_every=50, seed=22),
    # Metropolis, IP square with PBC (wrap), cutoff_range kernel, Zeeman, calc_r(Q), METROPOLIS_TIME
    "metro_ipsq_pbc": dict(lattice="IP_Square", a=1e-6, n=24, T=200, moment=2.37e-16, EB_kT=30, EB_dist="normal",
                           energies=[("dipolar", dict(cutoff_range=4e-6)), ("zeeman", dict(magnitude=2e-3, angle=1.0))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=8,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0, METROPOLIS_TIME=True),
                           kw=dict(PBC=True), update_kw=dict(Q=0.05), nsteps=150, snap_every=50, seed=23),
    # Metropolis, pinwheel, full kernel, sum branch with E_perp maintained
    "metro_pinwheel": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, moment=2.37e-16, EB_kT=60, EB_dist="lognormal",
                           energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=math.pi / 8))],
                           params=dict(UPDATE_SCHEME="Metropolis"), update_kw=dict(r=2.5e-6), nsteps=150, snap_every=50, seed=24),
    # Metropolis with the 'single' scheme: one uniformly drawn magnet per step (core.py:489-494)
    "metro_kagome_single": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16, EB_kT=40, EB_dist="normal",
                                energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.5))],
                                params=dict(UPDATE_SCHEME="Metropolis", MULTISAMPLING_SCHEME="single"),
                                update_kw={}, nsteps=300, snap_every=100, seed=25, count_singles=True),
    # grid selection whose supercells (R = 8 cells on a 10 x 10 grid) often all land outside the
    # grid, so _select_grid falls back to _select_single (core.py:555-558)
    "metro_oop_fallback": dict(lattice="OOP_Square", a=1e-6, n=10, T=300, moment=1.1e-15, EB_kT=20, EB_dist="normal",
                               energies=[("dipolar", {}), ("exchange", dict(J_kT=0.2))],
                               params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0),
                               update_kw=dict(r=8.5e-6), nsteps=150, snap_every=50, seed=26, count_singles=True),
    # progress() calls (core.py:736-760): Neel bound by MCsteps, by time (the last step gets the
    # remaining t_max - dt and does not switch, core.py:701-704, 757), by either, and a t_max
    # shorter than any switch (one no-switch attempt)
    "prog_neel_kagome": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16, EB_kT=20, EB_dist="normal",
                             energies=[("dipolar", {}), ("zeeman", dict(magnitude=5e-4, angle=0.3)), ("exchange", dict(J_kT=0.3))],
                             params=dict(UPDATE_SCHEME="Néel"),
                             calls=[dict(t_max=np.inf, MCsteps_max=0.5), dict(t_max=5e-8, MCsteps_max=np.inf),
                                    dict(t_max=2e-8, MCsteps_max=1.0), dict(t_max=1e-12, MCsteps_max=4.0),
                                    dict(t_max=1e-9, MCsteps_max=0.25), dict(t_max=3e-8, MCsteps_max=0.4)],
                             snap_every=2, seed=41),
    # Metropolis progress(MCsteps_max) (no time variable): whole and fractional sweeps
    "prog_metro_oop": dict(lattice="OOP_Square", a=1e-6, n=32, T=300, moment=1.1e-15, EB_kT=20, EB_dist="normal",
                           energies=[("dipolar", dict(cutoff_range=5e-6)), ("exchange", dict(J_kT=0.2))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=5,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3),
                           calls=[dict(MCsteps_max=0.5, r=4e-6), dict(MCsteps_max=1.25, r=4e-6),
                                  dict(t_max=np.inf, MCsteps_max=0.1, r=4e-6), dict(MCsteps_max=2.0, r=4e-6)],
                           snap_every=2, seed=42),
    # Metropolis with METROPOLIS_TIME: progress() bound by the summed sample times or by MCsteps
    "prog_metro_time": dict(lattice="IP_Square", a=1e-6, n=24, T=300, moment=2.37e-16, EB_kT=20, EB_dist="normal",
                            energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.8))],
                            params=dict(UPDATE_SCHEME="Metropolis", METROPOLIS_TIME=True),
                            calls=[dict(t_max=2e-9, MCsteps_max=np.inf, r=3e-6), dict(t_max=np.inf, MCsteps_max=0.5, r=3e-6),
                                   dict(t_max=1e-8, MCsteps_max=1.0, r=3e-6)],
                            snap_every=1, seed=43),
}


def main():
    core, energies, ASI = import_reference()
    only = set(sys.argv[1:])
    for name, spec in CASES.items():
        if only and name not in only:
            continue
        out = run_case(core, energies, ASI, spec)
        meta = {k: v for k, v in spec.items() if k not in ("energies", "params", "kw", "update_kw", "calls")}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), spec=np.array(repr(spec)), **out)
        print(f"{name}: {len(out['sw_y'])} switches, attempts {out['attempts'][-1]}, "
              f"t={out['t'][-1]:.4g}, E_tot {float(out['E_tot0']):.6g} -> {float(out['E_tot_final']):.6g}", meta["lattice"])


if __name__ == "__main__":
    main()
