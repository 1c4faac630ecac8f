"""Device-resident state of one Magnets instance and the calls into libhotspice_hip.so.

PyTorch is used only for device memory and the current HIP stream; all arithmetic on the
stepping path happens in the library's gfx950 kernels. Layout in HBM (row-major (ny, nx)):
  m int8 | flags uint8 | E, Ec, Ep, Epc f32 (combined fields + Kahan compensation), or E64 f64 |
  EB, beta, mu f32 | occ_list int32 | kernels K32/Kpo32 f32 and K64/Kps64/Kpo64 f64 of shape
  [n_sub][2kwy+1][2kwx+1] | run state hs_state_t[2] | scratch (Neel partials, Metropolis lists).
"""
import ctypes as C
import math

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


class DeviceState:
    def __init__(self, mm, device=None):
        torch = _torch()
        if not torch.cuda.is_available():
            raise _lib.HotspiceError("hotspice_amd needs a HIP device (MI355X): torch.cuda.is_available() is False")
        self.lib = _lib.load()
        self.torch = torch
        self.device = torch.device("cuda", 0 if device is None else int(device))
        ny, nx = mm.ny, mm.nx
        self.nx, self.ny = nx, ny
        n = nx * ny
        dev = self.device
        self.m = torch.zeros(n, dtype=torch.int8, device=dev)
        self.flags = torch.zeros(n, dtype=torch.uint8, device=dev)
        self.E = torch.zeros(n, dtype=torch.float32, device=dev)
        self.Ec = torch.zeros(n, dtype=torch.float32, device=dev)
        self.Ep = torch.zeros(n, dtype=torch.float32, device=dev)
        self.Epc = torch.zeros(n, dtype=torch.float32, device=dev)
        # the combined field in FP64 (Magnets(fp64_fields=True)); E/Ec are then unused
        self.E64 = torch.zeros(n, dtype=torch.float64, device=dev) if getattr(mm, "fp64_fields", False) else None
        self.EB = torch.zeros(n, dtype=torch.float32, device=dev)
        self.beta = torch.zeros(n, dtype=torch.float32, device=dev)
        self.beta64c = None  # FP64 per-cell beta, kept only while T is not uniform (Neel waiting times)
        self.mu = torch.zeros(n, dtype=torch.float32, device=dev)
        self.mu64 = torch.zeros(n, dtype=torch.float64, device=dev)
        self.pw64 = None  # FP64 pair weights when they are not the moments (DiMonopolarEnergy)
        self.src_ones = None  # source weights of DiMonopolarEnergy (ones)
        self.occ_list = torch.zeros(max(1, n), dtype=torch.int32, device=dev)
        self.state = torch.zeros(2 * _lib.STATE_BYTES, dtype=torch.uint8, device=dev)
        self.partials = torch.zeros(2 * 1024 * 32 + 4 * 1024, dtype=torch.uint8, device=dev)
        self.zee_bx = None
        self.zee_by = None
        self.kern = None  # dict of kernel tensors + window
        self.metro_ws = None
        self.relax_ws = None
        self.flip_ws = None
        self.wolff_ws = None
        self.min_cells = None
        self.min_result = None
        self.rec_scratch = None
        self.gen = 0  # bumped whenever the instance's parameters are refilled
        self._rec_need = None  # (gen, recompute scratch bytes)
        # FP64 shadow of the dipolar correlations (C64) and the spins it holds (m_ref): a full
        # recompute of every component stores it, a periodic recompute then refreshes the fields
        # from it by adding only the changed sources (hs_refresh_fields)
        self.C64 = None
        self.m_ref = None
        self._shadow_ver = 0  # bumped when kernels, weights or occupation change
        self._shadow = None  # (version, E_perp held) of the full recompute that stored the shadow
        self.shadow_refreshes = 0  # refreshes since that full recompute
        self.refreshes = 0  # refreshes run so far
        self.team_ws = None
        self.sel_cache = {}
        self.log_cap = 0
        self.log_idx = None
        self.log_dt = None
        self.inst = _lib.Instance()

    # ---- plumbing ----------------------------------------------------------------------
    # set to a list to collect (start event, end event, state pointers) around every team launch
    launch_events = None
    many_ws = None
    obs_scratch = None

    def stream(self):
        return C.c_void_p(self.torch.cuda.current_stream(self.device).cuda_stream)

    @staticmethod
    def ptr(t):
        return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(0)

    def upload(self, tensor, array):
        src = self.torch.from_numpy(np.ascontiguousarray(array).reshape(-1))
        tensor.copy_(src.to(tensor.dtype), non_blocking=False)

    def download(self, tensor):
        return tensor.cpu().numpy()

    # ---- run state ---------------------------------------------------------------------
    def read_state(self):
        raw = self.state[: _lib.STATE_BYTES].cpu().numpy().tobytes()
        return _lib.State.from_buffer_copy(raw)

    def write_state(self, st):
        buf = np.frombuffer(bytes(st), dtype=np.uint8)
        self.state[: _lib.STATE_BYTES].copy_(self.torch.from_numpy(buf.copy()))

    # ---- kernels -------------------------------------------------------------------------
    def build_kernels(self, geom: _lib.KernelGeom, rrx, rry, in_plane, mono=None):
        torch = self.torch
        dev = self.device
        kwx, kwy, ns = geom.kwx, geom.kwy, geom.n_sub
        shape = (ns, 2 * kwy + 1, 2 * kwx + 1)
        t_rrx = torch.from_numpy(np.ascontiguousarray(rrx, dtype=np.float64)).to(dev)
        t_rry = torch.from_numpy(np.ascontiguousarray(rry, dtype=np.float64)).to(dev)
        geom.rrx = t_rrx.data_ptr()
        geom.rry = t_rry.data_ptr()
        K64 = torch.empty(shape, dtype=torch.float64, device=dev)
        K32 = torch.empty(shape, dtype=torch.float32, device=dev)
        Kps64 = torch.empty(shape, dtype=torch.float64, device=dev) if in_plane else None
        Kpo64 = torch.empty(shape, dtype=torch.float64, device=dev) if in_plane else None
        Kpo32 = torch.empty(shape, dtype=torch.float32, device=dev) if in_plane else None
        if mono is not None:  # DiMonopolarEnergy (rrx / rry: absolute positions; geometry table on the device)
            raw = np.frombuffer(C.string_at(C.addressof(mono), C.sizeof(mono)), dtype=np.uint8).copy()
            t_mono = torch.from_numpy(raw).to(dev)
            _lib.check(self.lib.hs_build_dimonopolar_kernels(C.byref(geom), C.c_void_p(t_mono.data_ptr()), self.ptr(K64),
                                                             self.ptr(Kps64), self.ptr(Kpo64), self.ptr(K32),
                                                             self.ptr(Kpo32), self.stream()))
        else:
            _lib.check(self.lib.hs_build_dipolar_kernels(C.byref(geom), self.ptr(K64), self.ptr(Kps64),
                                                         self.ptr(Kpo64), self.ptr(K32), self.ptr(Kpo32),
                                                         self.stream()))
        torch.cuda.current_stream(dev).synchronize()
        self.invalidate_shadow()
        self.kern = dict(K64=K64, K32=K32, Kps64=Kps64, Kpo64=Kpo64, Kpo32=Kpo32, kwx=kwx, kwy=kwy, n_sub=ns,
                         _rr=(t_rrx, t_rry))

    def drop_kernels(self):
        self.kern = None

    def kernels_full(self, which, nx, ny):
        """Kernels zero-padded to the reference's (n_sub, 2ny-1, 2nx-1) shape, FP64 on the host."""
        if self.kern is None:
            return None
        name = {"K": "K64", "Kps": "Kps64", "Kpo": "Kpo64"}[which]
        t = self.kern[name]
        ns, kwx, kwy = self.kern["n_sub"], self.kern["kwx"], self.kern["kwy"]
        out = np.zeros((ns, 2 * ny - 1, 2 * nx - 1))
        if t is None:
            return out
        out[:, ny - 1 - kwy:ny + kwy, nx - 1 - kwx:nx + kwx] = t.cpu().numpy()
        return out

    # ---- FP64 shadow -------------------------------------------------------------------------
    # a refresh adds at most this many refreshes' rounding before the next full recompute, and is
    # taken only while the changed cells are at most this fraction of the grid (above it the full
    # correlation costs less than the list-driven additions)
    SHADOW_MAX_REFRESHES = 64
    SHADOW_MAX_CHANGED = 1 / 16

    def shadow_ptrs(self, in_plane, enabled=True):
        """Device pointers (C64, m_ref) for the instance descriptor, allocating them on first use
        (C64 holds E and E_perp correlations for in-plane lattices); NULLs when disabled."""
        if not enabled:
            self.C64 = self.m_ref = None
            self.invalidate_shadow()
            return self.ptr(None), self.ptr(None)
        n = self.nx * self.ny
        need = (2 if in_plane else 1) * n
        if self.C64 is None or self.C64.numel() < need:
            self.C64 = self.torch.zeros(need, dtype=self.torch.float64, device=self.device)
            self.m_ref = self.torch.zeros(n, dtype=self.torch.int8, device=self.device)
            self.invalidate_shadow()
        return self.ptr(self.C64), self.ptr(self.m_ref)

    def invalidate_shadow(self):
        """The next periodic recompute is a full one: the kernels, pair weights or occupation the
        shadow was computed with changed (the dipolar prefactor, Zeeman, exchange, temperatures
        and barriers do not enter it; the finishing pass applies them)."""
        self._shadow_ver += 1
        self._shadow = None

    def _stores_shadow(self, which):
        I = self.inst
        return bool(I.C64) and I.dip_pf != 0.0 and bool(which & 1) and bool(which & 2 or not I.use_perp)

    def shadow_usable(self, which, changed):
        """True when a refresh from the shadow reproduces the full recompute `which` for at most
        `changed` changed cells (None: unknown)."""
        return (changed is not None and changed <= self.SHADOW_MAX_CHANGED * self.nx * self.ny
                and self.nx <= 32767 and self.ny <= 32767  # hs_refresh_fields' limit
                and self._shadow == (self._shadow_ver, bool(self.inst.use_perp))
                and self.shadow_refreshes < self.SHADOW_MAX_REFRESHES and self._stores_shadow(which))

    def _shadow_stored(self, which):
        if self._stores_shadow(which):
            self._shadow = (self._shadow_ver, bool(self.inst.use_perp))
            self.shadow_refreshes = 0

    def refresh(self, which):
        """hs_refresh_fields: the fields of the current spins from the shadow."""
        scr = self.recompute_scratch()
        _lib.check(self.lib.hs_refresh_fields(C.byref(self.inst), which, self.ptr(scr), scr.numel(), self.stream()))
        self.shadow_refreshes += 1
        self.refreshes += 1

    # ---- calls -----------------------------------------------------------------------------
    def recompute(self, which=3, changed=None):
        """FP64 recompute of the fields `which`: a refresh from the shadow when `changed` (an upper
        bound on the cells whose spin changed since the last recompute) allows it, else a full one."""
        if self.shadow_usable(which, changed):
            self.refresh(which)
            return
        scr = self.recompute_scratch()
        _lib.check(self.lib.hs_recompute_fields(C.byref(self.inst), which, self.ptr(scr), scr.numel(), self.stream()))
        self._shadow_stored(which)

    def recompute_scratch(self):
        """The recompute scratch tensor of this instance, sized for it (the size is asked of the
        library once per refill of the instance's parameters)."""
        if self._rec_need is None or self._rec_need[0] != self.gen:
            need = int(self.lib.hs_recompute_scratch_bytes(C.byref(self.inst)))
            if need < 0:
                raise _lib.HotspiceError("bad instance for recompute")
            if self.inst.C64:
                need = max(need, int(self.lib.hs_refresh_scratch_bytes(C.byref(self.inst))))
            self._rec_need = (self.gen, need)
        need = self._rec_need[1]
        if self.rec_scratch is None or self.rec_scratch.numel() < need:
            self.rec_scratch = self.torch.empty(max(need, 8), dtype=self.torch.uint8, device=self.device)
        return self.rec_scratch

    def recompute_many(self, devs, which, changed=None):
        """hs_recompute_fields for several instances (DeviceStates on this device) in one call;
        those whose shadow allows it (changed[i]: as in recompute) are refreshed from it instead."""
        if changed is not None:
            keep = []
            for d, w, c in zip(devs, which, changed):
                if d.shadow_usable(w, c):
                    d.refresh(w)
                else:
                    keep.append((d, w))
            if not keep:
                return
            devs, which = [d for d, _ in keep], [w for _, w in keep]
        n = len(devs)
        insts = (C.c_void_p * n)(*[C.addressof(d.inst) for d in devs])
        wh = (C.c_int32 * n)(*which)
        scr = [d.recompute_scratch() for d in devs]
        ptrs = (C.c_void_p * n)(*[t.data_ptr() for t in scr])
        sizes = (C.c_int64 * n)(*[t.numel() for t in scr])
        need = int(self.lib.hs_recompute_many_ws_bytes(insts, wh, n))
        if need < 0:
            raise _lib.HotspiceError("bad instances for recompute")
        if self.many_ws is None or self.many_ws.numel() < need:
            self.many_ws = self.torch.empty(max(need, 256), dtype=self.torch.uint8, device=self.device)
        _lib.check(self.lib.hs_recompute_fields_many(insts, wh, ptrs, sizes, n, self.ptr(self.many_ws),
                                                     self.many_ws.numel(), self.stream()))
        for d, w in zip(devs, which):
            d._shadow_stored(w)

    def bins(self, x_start, y_start, weights=None):
        """A readout's bins and weights, uploaded once (see Bins)."""
        return Bins(self, x_start, y_start, weights)

    def binned_sums(self, bins):
        """hs_binned_sums over `bins` (from DeviceState.bins): the sums of m * w per bin, as a host
        array of shape (nw, nby, nbx)."""
        need = int(self.lib.hs_binned_sums_scratch_bytes(C.byref(self.inst), bins.nbx, bins.nw))
        if need < 0:
            raise _lib.HotspiceError("bad readout bins")
        if self.obs_scratch is None or self.obs_scratch.numel() < need:
            self.obs_scratch = self.torch.empty(max(need, 8), dtype=self.torch.uint8, device=self.device)
        out = self.torch.empty(bins.nw * bins.nby * bins.nbx, dtype=self.torch.float64, device=self.device)
        _lib.check(self.lib.hs_binned_sums(C.byref(self.inst), bins.nbx, bins.nby, self.ptr(bins.xs), self.ptr(bins.ys),
                                           self.ptr(bins.w), bins.nw, self.ptr(self.obs_scratch),
                                           self.obs_scratch.numel(), self.ptr(out), self.stream()))
        return out.cpu().numpy().reshape(bins.nw, bins.nby, bins.nbx)

    def components(self):
        out = self.torch.empty((6, self.ny, self.nx), dtype=self.torch.float64, device=self.device)
        _lib.check(self.lib.hs_field_components(C.byref(self.inst), self.ptr(out), self.stream()))
        return out.cpu().numpy()

    def energy_total(self):
        v = C.c_double(0.0)
        _lib.check(self.lib.hs_energy_total(C.byref(self.inst), C.byref(v), self.stream()))
        return v.value

    def zeeman_swap(self, Bx0, By0, bx_old=None, by_old=None):
        _lib.check(self.lib.hs_zeeman_swap(C.byref(self.inst), float(Bx0), float(By0), self.ptr(bx_old),
                                           self.ptr(by_old), self.stream()))

    def ensure_log(self, n):
        if self.log_cap < n:
            cap = max(n, 1024)
            self.log_idx = self.torch.zeros(cap, dtype=self.torch.int32, device=self.device)
            self.log_dt = self.torch.zeros(cap, dtype=self.torch.float64, device=self.device)
            self.log_cap = cap

    def neel_run(self, run: _lib.Run, persistent=True, nblocks=0):
        """Neel steps: one persistent team launch (default) or one launch per step."""
        if run.log:
            self.ensure_log(run.max_steps)
        if persistent:
            team = _lib.Team(inst=C.addressof(self.inst), sel=None, run=run, state=self.state.data_ptr(),
                             log_idx=self.log_idx.data_ptr() if run.log else None,
                             log_dt=self.log_dt.data_ptr() if run.log else None, log_cells=None, nblocks=int(nblocks))
            self.run_teams([team], "neel")
            return
        fc = C.c_int32(0)
        _lib.check(self.lib.hs_neel_run(C.byref(self.inst), C.byref(run), self.ptr(self.state), self.ptr(self.partials),
                                        self.ptr(self.log_idx) if run.log else C.c_void_p(0),
                                        self.ptr(self.log_dt) if run.log else C.c_void_p(0), C.byref(fc),
                                        self.stream()))

    def run_teams(self, teams, kind):
        """One cooperative launch stepping every team (instance) in `teams`: a list of _lib.Team or
        a NumPy array of the same records."""
        if isinstance(teams, np.ndarray):
            keep = teams
            arr = keep.ctypes.data_as(C.POINTER(_lib.Team))
            states = keep['state'].tolist()
        else:
            arr = (_lib.Team * len(teams))(*teams)
            states = [t.state for t in teams]
        lib = self.lib
        size_fn, run_fn = (lib.hs_neel_teams_ws_bytes, lib.hs_neel_run_teams) if kind == "neel" else \
            (lib.hs_metro_teams_ws_bytes, lib.hs_metro_run_teams)
        need = int(size_fn(arr, len(teams)))
        if need < 0:
            raise _lib.HotspiceError("bad team descriptors")
        if self.team_ws is None or self.team_ws.numel() < need:
            self.team_ws = self.torch.zeros(max(need, 256), dtype=self.torch.uint8, device=self.device)
        ev = self._event_start()
        _lib.check(run_fn(arr, len(teams), self.ptr(self.team_ws), self.team_ws.numel(), self.stream()))
        self._event_end(ev, states, kind)

    def _event_start(self):
        """Timing hook: when launch_events is a list, HIP events on the launch stream bracket the
        library call and (start, end, state pointers, kind) is appended to it. Setting
        DeviceState.launch_events (the class attribute) collects the launches of every instance."""
        if self.launch_events is None:
            return None
        e0 = self.torch.cuda.Event(enable_timing=True)
        e0.record(self.torch.cuda.current_stream(self.device))
        return e0

    def _event_end(self, e0, states, kind="metropolis"):
        if e0 is None:
            return
        e1 = self.torch.cuda.Event(enable_timing=True)
        e1.record(self.torch.cuda.current_stream(self.device))
        self.launch_events.append((e0, e1, states, kind))

    def metropolis_run(self, sel: _lib.Select, run: _lib.Run, log_cells=None, log_count=None):
        need = int(self.lib.hs_metropolis_ws_bytes(C.byref(self.inst), C.byref(sel)))
        if need < 0:
            raise _lib.HotspiceError("bad selection geometry")
        if self.metro_ws is None or self.metro_ws.numel() < need:
            self.metro_ws = self.torch.zeros(need, dtype=self.torch.uint8, device=self.device)
        ev = self._event_start()
        _lib.check(self.lib.hs_metropolis_run(C.byref(self.inst), C.byref(sel), C.byref(run), self.ptr(self.state),
                                              self.ptr(self.metro_ws), self.ptr(log_cells), self.ptr(log_count),
                                              self.stream()))
        self._event_end(ev, [self.state.data_ptr()])

    def relax_sweep(self, order, step_x, step_y, with_barrier, cutoff, reduced_k):
        """hs_relax_sweep: one Magnets._minimize_all over the subsets in `order`."""
        need = int(self.lib.hs_relax_ws_bytes(C.byref(self.inst), int(step_x), int(step_y)))
        if need < 0:
            raise _lib.HotspiceError("bad relax subsets")
        if self.relax_ws is None or self.relax_ws.numel() < need:
            self.relax_ws = self.torch.zeros(need, dtype=self.torch.uint8, device=self.device)
        o = np.ascontiguousarray(order, dtype=np.int32)
        _lib.check(self.lib.hs_relax_sweep(C.byref(self.inst), o.ctypes.data_as(C.c_void_p), int(o.size), int(step_x),
                                           int(step_y), int(bool(with_barrier)), int(cutoff), int(reduced_k),
                                           self.ptr(self.state), self.ptr(self.relax_ws), self.relax_ws.numel(),
                                           self.stream()))

    def minimize_scan(self, with_barrier):
        """hs_minimize_scan: (n, argmin cell, smallest barrier); the n cells below -kT are left in
        self.min_cells[:n] (row-major order)."""
        t = self.torch
        if self.min_cells is None:
            self.min_cells = t.empty(self.nx * self.ny, dtype=t.int32, device=self.device)
            self.min_result = t.empty(3, dtype=t.int64, device=self.device)
        _lib.check(self.lib.hs_minimize_scan(C.byref(self.inst), int(bool(with_barrier)), self.ptr(self.min_cells),
                                             self.ptr(self.min_result), self.stream()))
        n, amin, bits = (int(v) for v in self.min_result.cpu().numpy())
        return n, amin, float(np.array([bits], dtype=np.int64).view(np.float64)[0])

    def flip_cells(self, cells, cutoff, reduced_k):
        """hs_flip_cells: the given distinct flat cell indices switch together."""
        cells = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1)
        if cells.size == 0:
            return
        if cells.min() < 0 or cells.max() >= self.nx * self.ny:
            raise IndexError("cell index outside the grid")
        need = int(self.lib.hs_flip_ws_bytes(C.byref(self.inst)))
        if self.flip_ws is None or self.flip_ws.numel() < need:
            self.flip_ws = self.torch.zeros(need, dtype=self.torch.uint8, device=self.device)
        t_cells = self.torch.from_numpy(cells).to(self.device)
        _lib.check(self.lib.hs_flip_cells(C.byref(self.inst), self.ptr(t_cells), int(cells.size), int(cutoff),
                                          int(reduced_k), self.ptr(self.state), self.ptr(self.flip_ws),
                                          self.flip_ws.numel(), self.stream()))

    def wolff_run(self, run: _lib.Run, cutoff, reduced_k, log_cells=None, log_count=None):
        """hs_wolff_run: Wolff cluster steps until run's stop predicate or step budget."""
        need = int(self.lib.hs_wolff_ws_bytes(C.byref(self.inst)))
        if need < 0:
            raise _lib.HotspiceError("bad instance for Wolff steps")
        if self.wolff_ws is None or self.wolff_ws.numel() < need:
            self.wolff_ws = self.torch.zeros(need, dtype=self.torch.uint8, device=self.device)
        _lib.check(self.lib.hs_wolff_run(C.byref(self.inst), C.byref(run), int(cutoff), int(reduced_k),
                                         self.ptr(self.state), self.ptr(self.wolff_ws), self.wolff_ws.numel(),
                                         self.ptr(log_cells), self.ptr(log_count), self.stream()))

    def philox_uniform53(self, seed, step, tag, index0, n, open_zero=False):
        out = self.torch.empty(max(1, n), dtype=self.torch.float64, device=self.device)
        _lib.check(self.lib.hs_philox_uniform53(seed & 0xFFFFFFFFFFFFFFFF, step, tag, index0, n, int(open_zero),
                                                self.ptr(out), self.stream()))
        return out[:n].cpu().numpy()


def kernel_window(rrx, rry, nx, ny, cutoff, pbc):
    """Half widths (kwx, kwy) of the kernel region that can be nonzero under `cutoff`."""
    if pbc or not math.isfinite(cutoff):
        return nx - 1, ny - 1
    lim = cutoff * (1 + 1e-9)  # a margin: entries at the edge that round to zero are harmless
    cx = np.abs(np.asarray(rrx)) <= lim
    cy = np.abs(np.asarray(rry)) <= lim
    kwx = int(np.max(np.abs(np.nonzero(cx)[1] - (nx - 1)))) if cx.any() else 0
    kwy = int(np.max(np.abs(np.nonzero(cy)[1] - (ny - 1)))) if cy.any() else 0
    return min(kwx, nx - 1), min(kwy, ny - 1)


class Bins:
    """Bins and weights of a device readout: column bounds x_start (nbx+1 ascending entries from 0
    to nx), row bounds y_start (nby+1, 0 to ny) and optional FP64 weights (nw, ny, nx)."""

    def __init__(self, dev, x_start, y_start, weights=None):
        t = dev.torch
        self.nbx, self.nby = len(x_start) - 1, len(y_start) - 1
        self.xs = t.tensor(np.asarray(x_start, dtype=np.int32), device=dev.device)
        self.ys = t.tensor(np.asarray(y_start, dtype=np.int32), device=dev.device)
        if weights is None:
            self.w, self.nw = None, 1
        else:
            w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).reshape(-1, dev.ny, dev.nx))
            self.w, self.nw = t.tensor(w, device=dev.device), w.shape[0]
