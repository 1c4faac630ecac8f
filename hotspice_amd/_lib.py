"""ctypes bindings of libhotspice_hip.so (the C-ABI declared in include/hotspice_hip.h).

The library is loaded from this package's directory; it is built by ``hotspice_amd.build``
(``__graft_entry__.build()``). There is no fallback: if the library or a HIP device is missing,
every stepping call raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhotspice_hip.so")

HS_MAX_UC = 64
HS_MAX_NN = 49
HS_FLAG_OCC = 1
HS_FLAG_NEEL = 2

TAG_NEEL, TAG_SEL, TAG_CHOICE, TAG_ACCEPT, TAG_MTIME, TAG_WOLFF = 1, 2, 3, 4, 5, 7

EXPORTS = ("hs_version", "hs_last_error", "hs_abi_layout", "hs_device_info", "hs_philox_uniform53", "hs_build_dipolar_kernels",
           "hs_build_dimonopolar_kernels",
           "hs_recompute_fields", "hs_recompute_scratch_bytes", "hs_field_components", "hs_energy_total", "hs_zeeman_swap", "hs_neel_run",
           "hs_metropolis_run", "hs_metropolis_ws_bytes", "hs_neel_run_teams", "hs_neel_teams_ws_bytes",
           "hs_metro_run_teams", "hs_metro_teams_ws_bytes", "hs_recompute_fields_many", "hs_recompute_many_ws_bytes",
           "hs_binned_sums", "hs_binned_sums_scratch_bytes", "hs_relax_sweep", "hs_relax_ws_bytes",
           "hs_minimize_scan", "hs_flip_cells", "hs_flip_ws_bytes", "hs_wolff_run", "hs_wolff_ws_bytes",
           "hs_metropolis_run_static", "hs_set_graph_steps", "hs_graph_cache_clear", "hs_exchange_round_us", "hs_refresh_fields",
           "hs_refresh_scratch_bytes")

HS_METRO_STATIC_SCRATCH_BYTES = 4 << 20


class Instance(C.Structure):
    """hs_instance_t"""
    _fields_ = [
        ("m", C.c_void_p), ("flags", C.c_void_p), ("E", C.c_void_p), ("Ec", C.c_void_p),
        ("Ep", C.c_void_p), ("Epc", C.c_void_p), ("EB", C.c_void_p), ("beta", C.c_void_p), ("mu", C.c_void_p),
        ("mu_src", C.c_void_p), ("mu64", C.c_void_p), ("pw64", C.c_void_p), ("K32", C.c_void_p), ("Kpo32", C.c_void_p), ("K64", C.c_void_p), ("Kps64", C.c_void_p),
        ("zee_bx", C.c_void_p), ("zee_by", C.c_void_p), ("occ_list", C.c_void_p), ("beta64c", C.c_void_p),
        ("E64", C.c_void_p), ("C64", C.c_void_p), ("m_ref", C.c_void_p),
        ("seed", C.c_uint64), ("zee_Bx", C.c_double), ("zee_By", C.c_double), ("dip_pf", C.c_double),
        ("J", C.c_double), ("m_perp_factor", C.c_double), ("beta64", C.c_double),
        ("nx", C.c_int32), ("ny", C.c_int32), ("ux", C.c_int32), ("uy", C.c_int32), ("n_sub", C.c_int32),
        ("kwx", C.c_int32), ("kwy", C.c_int32), ("pbc", C.c_int32), ("in_plane", C.c_int32),
        ("use_perp", C.c_int32), ("barrier", C.c_int32), ("nn_w", C.c_int32), ("nn_h", C.c_int32),
        ("n_occ", C.c_int32), ("has_zeeman", C.c_int32), ("uniform_beta", C.c_int32), ("pair_src_unit", C.c_int32),
        ("sub_of", C.c_int32 * HS_MAX_UC), ("ox_uc", C.c_double * HS_MAX_UC), ("oy_uc", C.c_double * HS_MAX_UC),
        ("nn_mask", C.c_float * HS_MAX_NN),
    ]


class KernelGeom(C.Structure):
    """hs_kernel_geom_t"""
    _fields_ = [
        ("rrx", C.c_void_p), ("rry", C.c_void_p), ("I", C.c_double), ("cutoff", C.c_double), ("decay", C.c_double),
        ("nx", C.c_int32), ("ny", C.c_int32), ("ux", C.c_int32), ("uy", C.c_int32), ("n_sub", C.c_int32),
        ("kwx", C.c_int32), ("kwy", C.c_int32), ("pbc", C.c_int32), ("in_plane", C.c_int32),
        ("site_x", C.c_int32 * HS_MAX_UC), ("site_y", C.c_int32 * HS_MAX_UC),
        ("shift_x", C.c_int32 * HS_MAX_UC), ("shift_y", C.c_int32 * HS_MAX_UC),
        ("uc_ox", C.c_double * HS_MAX_UC), ("uc_oy", C.c_double * HS_MAX_UC), ("uc_o", C.c_double * HS_MAX_UC),
    ]


class MonopoleGeom(C.Structure):
    """hs_monopole_geom_t"""
    _fields_ = [("ca", C.c_double * HS_MAX_UC), ("sa", C.c_double * HS_MAX_UC), ("cp", C.c_double * HS_MAX_UC),
                ("sp", C.c_double * HS_MAX_UC), ("h", C.c_double * HS_MAX_UC), ("h_small", C.c_double)]


class Run(C.Structure):
    """hs_run_t"""
    _fields_ = [("t_max", C.c_double), ("mc_max", C.c_double), ("attempt_freq", C.c_double),
                ("max_steps", C.c_int64), ("log", C.c_int32), ("nblocks", C.c_int32), ("begin", C.c_int32),
                ("first_free", C.c_int32)]


class State(C.Structure):
    """hs_state_t"""
    _fields_ = [("t", C.c_double), ("t0", C.c_double), ("attempts", C.c_int64), ("attempts0", C.c_int64),
                ("switches", C.c_int64), ("step", C.c_int64), ("steps_done", C.c_int64), ("done", C.c_int32),
                ("err", C.c_int32)]


class Select(C.Structure):
    """hs_select_t"""
    _fields_ = [("pat_off", C.c_void_p), ("pat_cells", C.c_void_p), ("Rx", C.c_int32), ("Ry", C.c_int32),
                ("snx", C.c_int32), ("sny", C.c_int32), ("ncols", C.c_int32), ("nsc", C.c_int32),
                ("scheme", C.c_int32), ("cutoff", C.c_int32), ("reduced_k", C.c_int32),
                ("metropolis_time", C.c_int32), ("r_x", C.c_double), ("r_y", C.c_double)]


class Team(C.Structure):
    """hs_team_t"""
    _fields_ = [("inst", C.c_void_p), ("sel", C.c_void_p), ("run", Run), ("state", C.c_void_p),
                ("log_idx", C.c_void_p), ("log_dt", C.c_void_p), ("log_cells", C.c_void_p),
                ("nblocks", C.c_int32), ("pad", C.c_int32)]


STATE_BYTES = C.sizeof(State)

_lib = None


class HotspiceError(RuntimeError):
    pass


def load():
    """Loads the library (once). Raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HotspiceError(f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    # PyTorch-ROCm brings its own HIP runtime under the system runtime's soname (libamdhip64.so.7).
    # Loaded after torch, the library binds torch's copy. Loaded before it, the library would pull
    # in the system HIP and HSA runtimes and torch a second pair: two runtimes in one process, and
    # the one that initialises second finds no device.
    import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
    sig = {
        "hs_version": ([], C.c_int),
        "hs_abi_layout": ([C.POINTER(i64), i32], C.c_int),
        "hs_last_error": ([], C.c_char_p),
        "hs_device_info": ([C.POINTER(i32), C.c_char_p, i32], C.c_int),
        "hs_philox_uniform53": ([u64, u64, C.c_uint32, i64, i64, i32, vp, vp], C.c_int),
        "hs_build_dipolar_kernels": ([C.POINTER(KernelGeom), vp, vp, vp, vp, vp, vp], C.c_int),
        "hs_build_dimonopolar_kernels": ([C.POINTER(KernelGeom), vp, vp, vp, vp, vp, vp, vp], C.c_int),
        "hs_recompute_fields": ([C.POINTER(Instance), i32, vp, i64, vp], C.c_int),
        "hs_recompute_scratch_bytes": ([C.POINTER(Instance)], C.c_int64),
        "hs_refresh_fields": ([C.POINTER(Instance), i32, vp, i64, vp], C.c_int),
        "hs_refresh_scratch_bytes": ([C.POINTER(Instance)], C.c_int64),
        "hs_field_components": ([C.POINTER(Instance), vp, vp], C.c_int),
        "hs_energy_total": ([C.POINTER(Instance), C.POINTER(dbl), vp], C.c_int),
        "hs_zeeman_swap": ([C.POINTER(Instance), dbl, dbl, vp, vp, vp], C.c_int),
        "hs_neel_run": ([C.POINTER(Instance), C.POINTER(Run), vp, vp, vp, vp, C.POINTER(i32), vp], C.c_int),
        "hs_metropolis_run": ([C.POINTER(Instance), C.POINTER(Select), C.POINTER(Run), vp, vp, vp, vp, vp], C.c_int),
        "hs_metropolis_ws_bytes": ([C.POINTER(Instance), C.POINTER(Select)], C.c_int64),
        "hs_metropolis_run_static": ([C.POINTER(Instance), C.POINTER(Select), C.POINTER(Run), vp, vp, vp, vp], C.c_int),
        "hs_neel_run_teams": ([C.POINTER(Team), i32, vp, i64, vp], C.c_int),
        "hs_neel_teams_ws_bytes": ([C.POINTER(Team), i32], C.c_int64),
        "hs_metro_run_teams": ([C.POINTER(Team), i32, vp, i64, vp], C.c_int),
        "hs_set_graph_steps": ([i32], C.c_int),
        "hs_graph_cache_clear": ([], C.c_int),
        "hs_exchange_round_us": ([i32, i32, C.POINTER(C.c_double), C.POINTER(C.c_int32)], C.c_int),
        "hs_metro_teams_ws_bytes": ([C.POINTER(Team), i32], C.c_int64),
        "hs_recompute_fields_many": ([C.POINTER(vp), C.POINTER(i32), C.POINTER(vp), C.POINTER(i64), i32, vp, i64, vp],
                                     C.c_int),
        "hs_recompute_many_ws_bytes": ([C.POINTER(vp), C.POINTER(i32), i32], C.c_int64),
        "hs_binned_sums": ([C.POINTER(Instance), i32, i32, vp, vp, vp, i32, vp, i64, vp, vp], C.c_int),
        "hs_binned_sums_scratch_bytes": ([C.POINTER(Instance), i32, i32], C.c_int64),
        "hs_relax_sweep": ([C.POINTER(Instance), vp, i32, i32, i32, i32, i32, i32, vp, vp, i64, vp], C.c_int),
        "hs_relax_ws_bytes": ([C.POINTER(Instance), i32, i32], C.c_int64),
        "hs_minimize_scan": ([C.POINTER(Instance), i32, vp, vp, vp], C.c_int),
        "hs_flip_cells": ([C.POINTER(Instance), vp, i32, i32, i32, vp, vp, i64, vp], C.c_int),
        "hs_flip_ws_bytes": ([C.POINTER(Instance)], C.c_int64),
        "hs_wolff_run": ([C.POINTER(Instance), C.POINTER(Run), i32, i32, vp, vp, i64, vp, vp, vp], C.c_int),
        "hs_wolff_ws_bytes": ([C.POINTER(Instance)], C.c_int64),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _lib = lib
    return lib


def check(rc):
    if rc != 0:
        msg = load().hs_last_error().decode(errors="replace")
        raise HotspiceError(f"libhotspice_hip error {rc}: {msg}")
    return rc


def exported_symbols():
    """Names declared in include/hotspice_hip.h that the loaded library exports."""
    lib = load()
    return [n for n in EXPORTS if hasattr(lib, n)]


def abi_layout_python():
    """The same 20 numbers as hs_abi_layout, from the ctypes mirrors."""
    return [C.sizeof(Instance), Instance.seed.offset, Instance.nx.offset, Instance.sub_of.offset,
            Instance.nn_mask.offset, C.sizeof(KernelGeom), KernelGeom.nx.offset, KernelGeom.uc_o.offset,
            C.sizeof(Run), Run.max_steps.offset, Run.begin.offset, C.sizeof(State), State.done.offset,
            C.sizeof(Select), Select.metropolis_time.offset, C.sizeof(Team), Team.run.offset, Team.nblocks.offset,
            Instance.pair_src_unit.offset, C.sizeof(MonopoleGeom)]


def abi_layout_library():
    lib = load()
    out = (C.c_int64 * 20)()
    n = lib.hs_abi_layout(out, 20)
    return list(out)[:n]


def set_graph_steps(steps):
    """Steps of the batched Metropolis loop captured per hipGraph (0: plain launches); returns the
    previous setting. Results are bit-identical either way."""
    return int(load().hs_set_graph_steps(int(steps)))


def clear_graph_cache():
    """Destroys the cached step graphs (hs_graph_cache_clear); later graph-mode calls capture again."""
    check(load().hs_graph_cache_clear())


def exchange_round_us(G, rounds=20000):
    """Microseconds per cross-workgroup exchange round of G workgroups (the latency floor of one
    persistent Neel step with a team of G), and the rounds that timed out."""
    us, failed = C.c_double(0.0), C.c_int32(0)
    check(load().hs_exchange_round_us(int(G), int(rounds), C.byref(us), C.byref(failed)))
    return us.value, failed.value
