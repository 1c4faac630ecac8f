"""Shared builders for the parity tests: the golden-fixture specs, mirrored for the oracle
and for the product. Test infrastructure only."""
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
kB = 1.380649e-23

# Same specs as tests/golden/make_golden.py (kept in sync by test_golden_specs_match).
CASES = {
    "neel_pinwheel": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, moment=2.37e-16,
                          energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=math.pi / 8))],
                          params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=np.inf), nsteps=200, seed=11),
    "neel_oop_parabolic": dict(lattice="OOP_Square", a=1e-6, n=20, T=300, moment=1.1e-15,
                               energies=[("dipolar", {}), ("zeeman", dict(magnitude=2e-5)), ("exchange", dict(J_kT=0.5))],
                               params=dict(UPDATE_SCHEME="Néel", ENERGY_BARRIER_METHOD="parabolic"),
                               update_kw=dict(t_max=np.inf), nsteps=200, seed=12),
    "neel_kagome_tmax": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16,
                             energies=[("dipolar", {}), ("zeeman", dict(magnitude=5e-4, angle=0.3)), ("exchange", dict(J_kT=0.3))],
                             params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=1e-3), nsteps=120, seed=13),
    # a temperature that varies over the grid (T_ramp: T * (1 + T_ramp * (x / (nx-1) - 1/2)), the
    # array stored in the fixture): the grid's minimum barrier reorders the Neel waiting times
    # (core.py:686-699)
    "neel_pinwheel_tramp": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, T_ramp=0.6, moment=2.37e-16,
                                energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.4))],
                                params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=np.inf), nsteps=200, seed=14),
    "neel_oop_tramp_parabolic": dict(lattice="OOP_Square", a=1e-6, n=20, T=300, T_ramp=0.8, moment=1.1e-15,
                                     energies=[("dipolar", {}), ("exchange", dict(J_kT=0.3))],
                                     params=dict(UPDATE_SCHEME="Néel", ENERGY_BARRIER_METHOD="parabolic"),
                                     update_kw=dict(t_max=1e-3), nsteps=200, seed=15),
    "metro_oop_sum": dict(lattice="OOP_Square", a=1e-6, n=32, T=300, moment=1.1e-15,
                          energies=[("dipolar", {})], params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0),
                          update_kw=dict(r=4e-6), nsteps=150, seed=21),
    "metro_oop_conv": dict(lattice="OOP_Square", a=1e-6, n=32, T=600, moment=1.1e-15,
                           energies=[("dipolar", {}), ("exchange", dict(J_kT=0.2))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=5,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=2),
                           update_kw=dict(r=3e-6), nsteps=150, seed=22),
    "metro_ipsq_pbc": dict(lattice="IP_Square", a=1e-6, n=24, T=200, moment=2.37e-16,
                           energies=[("dipolar", dict(cutoff_range=4e-6)), ("zeeman", dict(magnitude=2e-3, angle=1.0))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=8,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0, METROPOLIS_TIME=True),
                           kw=dict(PBC=True), update_kw=dict(Q=0.05), nsteps=150, seed=23),
    "metro_pinwheel": dict(lattice="IP_Pinwheel", a=1e-6, n=24, T=300, moment=2.37e-16,
                           energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=math.pi / 8))],
                           params=dict(UPDATE_SCHEME="Metropolis"), update_kw=dict(r=2.5e-6), nsteps=150, seed=24),
    "metro_kagome_single": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16,
                                energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.5))],
                                params=dict(UPDATE_SCHEME="Metropolis", MULTISAMPLING_SCHEME="single"),
                                update_kw={}, nsteps=300, seed=25),
    "metro_oop_fallback": dict(lattice="OOP_Square", a=1e-6, n=10, T=300, moment=1.1e-15,
                               energies=[("dipolar", {}), ("exchange", dict(J_kT=0.2))],
                               params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=0),
                               update_kw=dict(r=8.5e-6), nsteps=150, seed=26),
}

# Reference fixtures of progress() calls (tests/golden/make_golden.py, "calls"): one entry per call.
PROGRESS = {
    "prog_neel_kagome": dict(lattice="IP_Kagome", a=1e-6, n=20, T=300, moment=2.37e-16,
                             energies=[("dipolar", {}), ("zeeman", dict(magnitude=5e-4, angle=0.3)), ("exchange", dict(J_kT=0.3))],
                             params=dict(UPDATE_SCHEME="Néel"),
                             calls=[dict(t_max=np.inf, MCsteps_max=0.5), dict(t_max=5e-8, MCsteps_max=np.inf),
                                    dict(t_max=2e-8, MCsteps_max=1.0), dict(t_max=1e-12, MCsteps_max=4.0),
                                    dict(t_max=1e-9, MCsteps_max=0.25), dict(t_max=3e-8, MCsteps_max=0.4)], seed=41),
    "prog_metro_oop": dict(lattice="OOP_Square", a=1e-6, n=32, T=300, moment=1.1e-15,
                           energies=[("dipolar", dict(cutoff_range=5e-6)), ("exchange", dict(J_kT=0.2))],
                           params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=5,
                                       SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=3),
                           calls=[dict(MCsteps_max=0.5, r=4e-6), dict(MCsteps_max=1.25, r=4e-6),
                                  dict(t_max=np.inf, MCsteps_max=0.1, r=4e-6), dict(MCsteps_max=2.0, r=4e-6)], seed=42),
    "prog_metro_time": dict(lattice="IP_Square", a=1e-6, n=24, T=300, moment=2.37e-16,
                            energies=[("dipolar", {}), ("zeeman", dict(magnitude=1e-3, angle=0.8))],
                            params=dict(UPDATE_SCHEME="Metropolis", METROPOLIS_TIME=True),
                            calls=[dict(t_max=2e-9, MCsteps_max=np.inf, r=3e-6), dict(t_max=np.inf, MCsteps_max=0.5, r=3e-6),
                                   dict(t_max=1e-8, MCsteps_max=1.0, r=3e-6)], seed=43),
}


# Oracle-only cases (no reference fixture): seeded synthetic pattern and E_B. The window case
# has a short dipolar cutoff with a matching REDUCED_KERNEL_SIZE and no PBC, so the Metropolis
# update reaches its sources both by supercell lookup and by filtering the whole list.
SYNTH = {
    "metro_oop_window": dict(lattice="OOP_Square", a=1e-6, n=64, T=900, moment=1.1e-15,
                             energies=[("dipolar", dict(cutoff_range=3e-6)), ("exchange", dict(J_kT=0.1))],
                             params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=3,
                                         SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0),
                             update_kw=dict(r=6e-6), nsteps=150, seed=31),
    # above the persistent-kernel size (80 x 80 cells): the two-launch step
    "metro_oop_window_large": dict(lattice="OOP_Square", a=1e-6, n=136, T=900, moment=1.1e-15,
                                   energies=[("dipolar", dict(cutoff_range=3e-6)), ("exchange", dict(J_kT=0.1))],
                                   params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=3,
                                               SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0),
                                   update_kw=dict(r=6e-6), nsteps=60, seed=33),
    # two-launch step with PBC: truncated windows wrap, in-plane exchange stencil wraps
    "metro_ipsq_pbc_large": dict(lattice="IP_Square", a=1e-6, n=96, T=300, moment=2.37e-16,
                                 energies=[("dipolar", dict(cutoff_range=4e-6)), ("exchange", dict(J_kT=0.1)),
                                           ("zeeman", dict(magnitude=1e-3, angle=0.4))],
                                 params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=8,
                                             SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0),
                                 kw=dict(PBC=True), update_kw=dict(r=10e-6), nsteps=60, seed=34),
    # DiMonopolarEnergy: the same stepping kernels with dumbbell kernels and charge weights
    "neel_kagome_mono": dict(lattice="IP_Kagome", a=1e-6, n=40, T=300, moment=2.37e-16,
                             energies=[("dimonopolar", dict(d=350e-9, small_d=150e-9)),
                                       ("zeeman", dict(magnitude=5e-4, angle=0.3))],
                             params=dict(UPDATE_SCHEME="Néel"), update_kw=dict(t_max=np.inf), nsteps=150, seed=35),
    "metro_ipsq_mono_pbc": dict(lattice="IP_Square", a=1e-6, n=24, T=300, moment=2.37e-16,
                                energies=[("dimonopolar", dict(d=400e-9, small_d=200e-9)),
                                          ("zeeman", dict(magnitude=1e-3, angle=0.4))],
                                params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=8,
                                            SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=2),
                                kw=dict(PBC=True), update_kw=dict(r=2e-6), nsteps=120, seed=36),
    # more samples than selection threads: the presampled step (wide sampling launch; tiles
    # found by a wide launch without PBC, by the selection workgroup with PBC)
    "metro_oop_presample": dict(lattice="OOP_Square", a=1e-6, n=200, T=900, moment=1.1e-15,
                                energies=[("dipolar", dict(cutoff_range=3e-6)), ("exchange", dict(J_kT=0.1))],
                                params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=3,
                                            SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0),
                                update_kw=dict(r=2e-6), nsteps=40, seed=37),
    "metro_ipsq_presample_pbc": dict(lattice="IP_Square", a=1e-6, n=192, T=300, moment=2.37e-16,
                                     energies=[("dipolar", dict(cutoff_range=2e-6)),
                                               ("zeeman", dict(magnitude=1e-3, angle=0.4))],
                                     params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=4,
                                                 SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=0),
                                     kw=dict(PBC=True), update_kw=dict(r=1e-6), nsteps=30, seed=38),
    "metro_ipsq_window": dict(lattice="IP_Square", a=1e-6, n=48, T=400, moment=2.37e-16,
                              energies=[("dipolar", dict(cutoff_range=3e-6)), ("zeeman", dict(magnitude=1e-3, angle=0.7))],
                              params=dict(UPDATE_SCHEME="Metropolis", REDUCED_KERNEL_SIZE=6,
                                          SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=1),
                              update_kw=dict(r=5e-6), nsteps=150, seed=32),
}


def spec(name):
    for table in (CASES, PROGRESS, SYNTH):
        if name in table:
            return table[name]
    raise KeyError(name)


def load(name):
    if name in SYNTH:
        from oracle import lattices
        sp = SYNTH[name]
        rng = np.random.default_rng(sp["seed"])
        g = lattices.make(sp["lattice"], sp["a"], n=sp["n"], PBC=sp.get("kw", {}).get("PBC", False))
        shape = (g.ny, g.nx)
        kT = kB * sp["T"]
        fx = dict(pattern=np.where(rng.random(shape) < 0.5, -1.0, 1.0),
                  E_B=20 * kT * rng.normal(1, 0.05, size=shape))
        if "T_ramp" in sp:
            ramp = np.linspace(-0.5, 0.5, shape[1])[None, :] * np.ones((shape[0], 1))
            fx["T"] = sp["T"] * (1 + sp["T_ramp"] * ramp)
        return fx
    return dict(np.load(os.path.join(GOLDEN, f"{name}.npz"), allow_pickle=False))


def oracle_magnets(name, fx=None):
    """OracleMagnets for a fixture case, with the fixture's pattern and E_B."""
    from oracle import model
    sp = spec(name)
    fx = load(name) if fx is None else fx
    kT = kB * sp["T"]
    ens = []
    for kind, kw in sp["energies"]:
        if kind == "dipolar":
            ens.append(model.Dipolar(**kw))
        elif kind == "dimonopolar":
            ens.append(model.DiMonopolar(**kw))
        elif kind == "zeeman":
            ens.append(model.Zeeman(**kw))
        else:
            ens.append(model.Exchange(J=kw["J_kT"] * kT))
    params = dict(sp["params"])
    params["UPDATE_SCHEME"] = "Neel" if params["UPDATE_SCHEME"] in ("Néel", "Neel") else params["UPDATE_SCHEME"]
    return model.make_magnets(sp["lattice"], sp["a"], n=sp["n"], T=fx.get("T", sp["T"]), E_B=fx["E_B"],
                              moment=sp["moment"], pattern=fx["pattern"], energies=ens, params=params,
                              seed=sp["seed"], PBC=sp.get("kw", {}).get("PBC", False))


def product_magnets(name, fx=None, **extra):
    """hotspice_amd Magnets for a fixture case (same pattern, E_B and Philox seed)."""
    import hotspice_amd as hs
    sp = spec(name)
    fx = load(name) if fx is None else fx
    kT = kB * sp["T"]
    ens = []
    for kind, kw in sp["energies"]:
        if kind == "dipolar":
            ens.append(hs.DipolarEnergy(**kw))
        elif kind == "dimonopolar":
            ens.append(hs.energies.DiMonopolarEnergy(**kw))
        elif kind == "zeeman":
            ens.append(hs.ZeemanEnergy(**kw))
        else:
            ens.append(hs.ExchangeEnergy(J=kw["J_kT"] * kT))
    cls = getattr(hs.ASI, sp["lattice"])
    return cls(sp["a"], sp["n"], T=fx.get("T", sp["T"]), E_B=fx["E_B"], moment=sp["moment"], pattern=fx["pattern"],
               energies=ens, params=hs.SimParams(**sp["params"]), seed=sp["seed"], **sp.get("kw", {}), **extra)


def energy_scale(om):
    """Magnitude that energy errors are measured against: the largest |E|, or the size of the
    individual terms where they cancel. Exchange between perpendicular neighbours (all of the
    pinwheel's) and a Zeeman field perpendicular to every magnet sum to zero up to the rounding of
    the orientations, which the FP32 path holds to ~1e-8 of the term, not of the zero."""
    from oracle import model
    scale = float(np.abs(om.E).max())
    for e in om._energies:
        if isinstance(e, model.Exchange):
            scale = max(scale, abs(e.J) * float(np.abs(om.nn).sum()))
        elif isinstance(e, model.Zeeman):
            scale = max(scale, float(np.max(np.abs(e._magnitude) * np.abs(om.moment))))
    return scale
