"""FP64 NumPy restatement of Hotspice's Monte Carlo stepping path (the oracle).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the cpu_baseline
leg of bench.py, as the checker. The product (hotspice_amd) never imports it.

Everything is float64 on (ny, nx) grids, as in the reference. Randomness is injectable:
  rng="philox": the keyed Philox streams of oracle/philox.py (the streams the HIP kernels
                draw), so trajectories can be replayed bit-for-bit;
  rng="numpy":  NumPy generators with the same call structure as the reference, seeded per
                instance (a Generator for the continuous draws, a RandomState for the legacy
                randint/choice calls the reference makes on the global np.random), for the
                CPU timing and the statistical comparisons.

Restated reference code (file:line in bvwaeyen/Hotspice):
  Energy contract ............ hotspice/energies.py:22-83
  ZeemanEnergy ............... hotspice/energies.py:86-140
  DipolarEnergy kernels ...... hotspice/energies.py:176-288
  DipolarEnergy.update ....... hotspice/energies.py:290-308
  DiMonopolarEnergy .......... hotspice/energies.py:394-751
  update_single .............. hotspice/energies.py:310-329
  update_multiple ............ hotspice/energies.py:331-366
  ExchangeEnergy ............. hotspice/energies.py:754-791
  update_energy/switch_energy  hotspice/core.py:241-271
  E_barrier .................. hotspice/core.py:625-656
  _update_Metropolis ......... hotspice/core.py:659-677
  _update_Neel ............... hotspice/core.py:679-709
  select / _select_grid ...... hotspice/core.py:456-558
  calc_r ..................... hotspice/core.py:720-733
  progress ................... hotspice/core.py:736-760, lower_than utils.py:90-92
  minimize / _minimize_all / relax  hotspice/core.py:763-845
  _update_Wolff / _select_cluster_exchange  hotspice/core.py:588-602, 711-718
"""
import math

import numpy as np
from scipy import signal

from . import philox
from . import lattices

kB = 1.380649e-23


def as_2D_array(value, shape):
    """utils.py:208-228"""
    arr = np.asarray(value, dtype=float)
    if arr.size == 1:
        return np.ones(shape, dtype=float) * float(arr.reshape(-1)[0])
    if arr.size != math.prod(shape):
        raise ValueError(f"Incorrect shape: passed array of shape {arr.shape} is not of desired shape {shape}.")
    return arr.reshape(shape).astype(float)


def lower_than(x, y, rtol=1e-10):
    """utils.py:90-92: False if x is within rtol of y or above it."""
    return x < y and not np.isclose(x, y, atol=0, rtol=rtol)


# --------------------------------------------------------------------------------------------
# Energies
# --------------------------------------------------------------------------------------------
class Energy:
    def initialize(self, mm):
        self.mm = mm
        self.E = np.zeros_like(mm.m)
        self.E_perp = np.zeros_like(mm.m)
        self._initialize()
        self.update()

    def energy_switch(self, idx=None):
        return -2 * self.E if idx is None else -2 * self.E[idx[0], idx[1]]


class Zeeman(Energy):
    """energies.py:86-140"""

    def __init__(self, magnitude=0.0, angle=0.0):
        self._magnitude = magnitude
        self._angle = angle

    def _initialize(self):
        self.set_field(self._magnitude, self._angle)

    def set_field(self, magnitude=None, angle=None):
        if magnitude is not None:
            self._magnitude = magnitude
        if angle is not None:
            self._angle = angle
        if not hasattr(self, "mm"):
            return
        mm = self.mm
        mag = as_2D_array(self._magnitude, mm.shape)
        ang = as_2D_array(self._angle, mm.shape)
        if mm.in_plane:
            bx, by = mag * np.cos(ang), mag * np.sin(ang)
            self.E_factor = -mm.moment * (bx * mm.orientation[:, :, 0] + by * mm.orientation[:, :, 1])
            self.E_factor_perp = -mm.moment * (by * mm.orientation[:, :, 0] - bx * mm.orientation[:, :, 1])
        else:
            self.E_factor = -mm.moment * np.copy(mag)
            self.E_factor_perp = np.zeros_like(mm.m)
        self.update()

    def update(self):
        self.E = self.mm.m * self.E_factor
        if self.mm.USE_PERP_ENERGY:
            self.E_perp = self.mm.m * self.E_factor_perp

    def update_single(self, idx):
        self.update_multiple(idx)

    def update_multiple(self, idx):
        self.E[idx[0], idx[1]] *= -1
        if self.mm.USE_PERP_ENERGY:
            self.E_perp[idx[0], idx[1]] *= -1

    @property
    def E_tot(self):
        return np.sum(self.E)


class Exchange(Energy):
    """energies.py:754-791 (incremental updates are full recomputes, as in the reference)."""

    def __init__(self, J=1.0):
        self.J = J

    def _initialize(self):
        self.local_interaction = self.mm.nn

    def update(self):
        mm = self.mm
        bnd = "wrap" if mm.PBC else "fill"
        if mm.in_plane:
            mx = mm.orientation[:, :, 0] * mm.m
            my = mm.orientation[:, :, 1] * mm.m
            smx = signal.convolve2d(mx, self.local_interaction, mode="same", boundary=bnd)
            smy = signal.convolve2d(my, self.local_interaction, mode="same", boundary=bnd)
            self.E = -self.J * (smx * mx + smy * my)
            if mm.USE_PERP_ENERGY:
                self.E_perp = -self.J * (smy * mx - smx * my)
        else:
            self.E = -self.J * signal.convolve2d(mm.m, self.local_interaction, mode="same", boundary=bnd) * mm.m
            if mm.USE_PERP_ENERGY:
                self.E_perp = np.zeros_like(mm.m)

    def update_single(self, idx):
        self.update()

    def update_multiple(self, idx):
        self.update()

    @property
    def E_tot(self):
        return np.sum(self.E) / 2


class Dipolar(Energy):
    """energies.py:143-391"""

    def __init__(self, prefactor=1.0, decay_exponent=-3, cutoff_range=np.inf, recompute="auto"):
        self.prefactor = prefactor
        self.decay_exponent = decay_exponent
        self.cutoff_range = cutoff_range
        self.recompute = recompute

    def _initialize(self):
        K, Kps, Kpo, idx = build_kernels(self.mm, self.decay_exponent, self.cutoff_range)
        self.kernel_unitcell, self.kernel_perpself_unitcell, self.kernel_perpother_unitcell = K, Kps, Kpo
        self.kernel_unitcell_indices = idx

    def _conv_valid(self, kernel_flipped, mmoment):
        mode = self.recompute
        if mode == "auto":
            mode = "fft" if kernel_flipped.size * mmoment.size > 4e8 else "direct"
        if mode == "fft":
            return signal.fftconvolve(kernel_flipped, mmoment, mode="valid")
        return signal.convolve2d(kernel_flipped, mmoment, mode="valid")

    def update(self):
        mm = self.mm
        ux, uy = mm.ux, mm.uy
        total = np.zeros_like(mm.m)
        total_perp = np.zeros_like(mm.m)
        mmoment = mm.m * self._w_src()
        for y in range(uy):
            for x in range(ux):
                n = self.kernel_unitcell_indices[y, x]
                if n < 0:
                    continue
                partial_m = np.zeros_like(mm.m)
                partial_m[y::uy, x::ux] = mm.m[y::uy, x::ux]
                total += partial_m * self._conv_valid(self.kernel_unitcell[n, ::-1, ::-1], mmoment)
                if mm.USE_PERP_ENERGY:
                    total_perp += partial_m * self._conv_valid(self.kernel_perpself_unitcell[n, ::-1, ::-1], mmoment)
        self.E = self.prefactor * self._w_rx() * total
        if mm.USE_PERP_ENERGY:
            self.E_perp = self.prefactor * self._w_rx() * total_perp

    def update_single(self, idx):
        mm = self.mm
        y, x = int(idx[0][0]), int(idx[1][0])
        n = self.kernel_unitcell_indices[y % mm.uy, x % mm.ux]
        if n < 0:
            return
        sl = np.s_[mm.ny - 1 - y:2 * mm.ny - 1 - y, mm.nx - 1 - x:2 * mm.nx - 1 - x]
        msrc, mrx = mm.m * self._w_src(), mm.m * self._w_rx()
        self.E += 2 * (self.prefactor * msrc[y, x] * np.multiply(mrx, self.kernel_unitcell[n][sl]))
        self.E[y, x] *= -1
        if mm.USE_PERP_ENERGY:
            self.E_perp += 2 * (self.prefactor * msrc[y, x] * np.multiply(mrx, self.kernel_perpother_unitcell[n][sl]))
            self.E_perp[y, x] *= -1

    def update_multiple(self, idx):
        mm = self.mm
        idx = np.asarray(idx)
        self.E[idx[0], idx[1]] *= -1
        mmoment = mm.m * self._w_src()
        mrx = mm.m * self._w_rx()
        if mm.USE_PERP_ENERGY:
            self.E_perp[idx[0], idx[1]] *= -1
        raveled = (idx[1] % mm.ux) + (idx[0] % mm.uy) * mm.ux
        binned = np.bincount(raveled)
        for i in binned.nonzero()[0]:
            yu, xu = divmod(int(i), mm.ux)
            n = self.kernel_unitcell_indices[yu, xu]
            if n < 0:
                continue
            kernel = self.kernel_unitcell[n]
            kernel_perp = self.kernel_perpother_unitcell[n] if mm.USE_PERP_ENERGY else None
            here = idx[:, raveled == i]
            if here.shape[1] > mm.params["SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF"]:
                sw = np.zeros_like(mm.m)
                sw[here[0], here[1]] = mmoment[here[0], here[1]]
                k = mm.params["REDUCED_KERNEL_SIZE"]
                kx, ky = min(k, mm.nx - 1), min(k, mm.ny - 1)
                win = np.s_[mm.ny - 1 - ky:mm.ny + ky, mm.nx - 1 - kx:mm.nx + kx]
                bnd = "wrap" if mm.PBC else "fill"
                conv = signal.convolve2d(sw, kernel[win] if k else kernel, mode="same", boundary=bnd)
                if mm.USE_PERP_ENERGY:
                    conv_perp = signal.convolve2d(sw, kernel_perp[win] if k else kernel_perp, mode="same", boundary=bnd)
            else:
                conv = np.zeros_like(mm.m)
                conv_perp = np.zeros_like(mm.m) if mm.USE_PERP_ENERGY else None
                for j in range(here.shape[1]):
                    y, x = here[0, j], here[1, j]
                    sl = np.s_[mm.ny - 1 - y:2 * mm.ny - 1 - y, mm.nx - 1 - x:2 * mm.nx - 1 - x]
                    conv += mmoment[y, x] * kernel[sl]
                    if mm.USE_PERP_ENERGY:
                        conv_perp += mmoment[y, x] * kernel_perp[sl]
            self.E += 2 * (self.prefactor * np.multiply(mrx, conv))
            if mm.USE_PERP_ENERGY:
                self.E_perp += 2 * (self.prefactor * np.multiply(mrx, conv_perp))

    def _w_src(self):
        """Weight of the source magnet of a pair term: its moment (energies.py:318, 341)."""
        return self.mm.moment

    def _w_rx(self):
        """Weight of the receiving magnet of a pair term: its moment (energies.py:303, 318)."""
        return self.mm.moment

    @property
    def E_tot(self):
        return np.sum(self.E) / 2


class DiMonopolar(Dipolar):
    """energies.py:394-751: dumbbells of charges q = moment/d. A pair term is
    prefactor * K * m_source * m_receiver * q_receiver^2 (energies.py:654-659, 669-671, 707-710)."""

    def __init__(self, prefactor=1.0, d=200e-9, small_d=200e-9, recompute="auto"):
        self.prefactor = prefactor
        self.d = d
        self.small_d = small_d
        self.recompute = recompute

    def _initialize(self):
        mm = self.mm
        dist_uc = np.asarray(self.d, dtype=float) if isinstance(self.d, np.ndarray) else as_2D_array(self.d, (mm.uy, mm.ux))
        self.dist_unitcell = dist_uc
        dist = np.tile(dist_uc, (math.ceil(mm.ny / mm.uy), math.ceil(mm.nx / mm.ux)))
        self.dist_exact = dist[:mm.ny, :mm.nx]
        self.charge = mm.moment / self.dist_exact
        self.charge_sq = self.charge ** 2
        K, Kps, Kpo, idx = build_monopole_kernels(mm, dist_uc, self.small_d)
        self.kernel_unitcell, self.kernel_perpself_unitcell, self.kernel_perpother_unitcell = K, Kps, Kpo
        self.kernel_unitcell_indices = idx

    def _w_src(self):
        return np.ones_like(self.mm.m)

    def _w_rx(self):
        return self.charge_sq


def _large_positions(spacing, n, u):
    """Positions of the enlarged grid's indices -(n-1) .. n+u-1 with index n-1 at 0 (energies.py:416-428)."""
    idx = np.arange(-(n - 1), n + u)
    steps = np.asarray(spacing)[idx % n]
    pos = np.zeros(idx.size)
    pos[1:] = np.cumsum(steps)[:-1]
    return pos - pos[n - 1]


def _inv_dist(dx, dy, f32):
    """|r|^-1 with 0 where the charges coincide. The plain kernel rounds the squares to FP32 and
    takes their FP32 power -0.5 (energies.py:508-525). That power is taken here as the correctly
    rounded FP32 value of sq^-0.5, formed in FP64 (two FP64 roundings stay far below half an FP32
    ulp): the last bit of the reference's own FP32 power depends on its array library, and NumPy's
    float32 pow misrounds about 1 in 5 of these values, which moves a net dumbbell kernel entry, a
    near-cancelling sum of four such terms, by up to ~1.5e-5 relative."""
    sq = dx ** 2 + dy ** 2
    if f32:
        sq32 = sq.astype(np.float32).astype(np.float64)
        out = np.zeros(sq.shape, dtype=np.float32)
        nz = sq32 != 0
        out[nz] = (1.0 / np.sqrt(sq32[nz])).astype(np.float32)
        return out
    sq[sq == 0] = np.inf
    return sq ** -0.5


def build_monopole_kernels(mm, dist_uc, small_d):
    """DiMonopolarEnergy._initialize kernels (energies.py:416-641): K, K_perpself, K_perpother of
    shape (n_sub, 2ny-1, 2nx-1) and the (uy, ux) site->kernel index map."""
    ux, uy, nx, ny = mm.ux, mm.uy, mm.nx, mm.ny
    xl, yl = _large_positions(mm.dx, nx, ux), _large_positions(mm.dy, ny, uy)
    XX, YY = np.meshgrid(xl, yl)
    iy = (np.arange(yl.size) - (ny - 1)) % uy
    ix = (np.arange(xl.size) - (nx - 1)) % ux
    ang_uc = np.where(mm.occupation[:uy, :ux] != 0, mm.geom.angles[:uy, :ux], 0.0)
    ang = ang_uc[np.ix_(iy, ix)]
    half = (dist_uc / 2)[np.ix_(iy, ix)]
    occ = mm.occupation[:uy, :ux][np.ix_(iy, ix)].astype(float)
    # charges (+ then -) of every magnet as it is, and turned by +90 degrees with small_d
    cx = (XX + half * np.cos(ang), XX - half * np.cos(ang))
    cy = (YY + half * np.sin(ang), YY - half * np.sin(ang))
    px = (XX + small_d / 2 * np.cos(ang + np.pi / 2), XX - small_d / 2 * np.cos(ang + np.pi / 2))
    py = (YY + small_d / 2 * np.sin(ang + np.pi / 2), YY - small_d / 2 * np.sin(ang + np.pi / 2))
    index = -np.ones((uy, ux), dtype=int)
    Ks, Kps, Kpo = [], [], []
    for y0 in range(uy):
        for x0 in range(ux):
            if mm.occupation[y0, x0] == 0:
                continue
            Yc, Xc = ny - 1 + y0, nx - 1 + x0
            sl = np.s_[y0:y0 + 2 * ny - 1, x0:x0 + 2 * nx - 1]
            cen = (ny - 1, nx - 1)

            def combine(tx, ty, sx, sy, f32):
                """NN + SS - NS - SN between targets (tx, ty) and the centre's charges (sx, sy)."""
                terms = []
                for a, b in ((0, 0), (1, 1), (0, 1), (1, 0)):
                    dx = tx[a][sl] - sx[b]
                    dy = ty[a][sl] - sy[b]
                    if a != b:  # no cross term of the centre with itself (energies.py:494-498)
                        dx[cen] = 0
                        dy[cen] = 0
                    terms.append(_inv_dist(dx, dy, f32) * occ[sl])
                return terms[0] + terms[1] + -terms[2] + -terms[3]

            c_n = ((cx[0][Yc, Xc], cx[1][Yc, Xc]), (cy[0][Yc, Xc], cy[1][Yc, Xc]))
            c_p = ((px[0][Yc, Xc], px[1][Yc, Xc]), (py[0][Yc, Xc], py[1][Yc, Xc]))
            kernel = combine(cx, cy, c_n[0], c_n[1], True)
            # perp-self: the centre turned; at the centre itself every term vanishes
            tx_s = tuple(a.copy() for a in cx)
            ty_s = tuple(a.copy() for a in cy)
            for q in range(2):
                tx_s[q][Yc, Xc], ty_s[q][Yc, Xc] = c_p[0][q], c_p[1][q]
            kernel_ps = combine(tx_s, ty_s, c_p[0], c_p[1], False)
            # perp-other: every other magnet turned, the centre as it is
            tx_o = tuple(a.copy() for a in px)
            ty_o = tuple(a.copy() for a in py)
            for q in range(2):
                tx_o[q][Yc, Xc], ty_o[q][Yc, Xc] = c_n[0][q], c_n[1][q]
            kernel_po = combine(tx_o, ty_o, c_n[0], c_n[1], False)
            if mm.PBC:
                for k in (kernel, kernel_ps, kernel_po):
                    _apply_pbc(k, nx, ny)
            index[y0, x0] = len(Ks)
            Ks.append(kernel * 1e-7)
            Kps.append(kernel_ps * 1e-7)
            Kpo.append(kernel_po * 1e-7)
    return np.asarray(Ks), np.asarray(Kps), np.asarray(Kpo), index


def build_kernels(mm, decay_exponent=-3, cutoff_range=np.inf):
    """DipolarEnergy._initialize (energies.py:176-288): per unit-cell site K, K_perpself, K_perpother.

    Returns arrays of shape (n_sub, 2ny-1, 2nx-1) and the (uy, ux) site->kernel index map.
    """
    ux, uy, nx, ny = mm.ux, mm.uy, mm.nx, mm.ny
    I = (mm.major_axis ** 2 + mm.minor_axis ** 2) / 16
    ncx = 2 * math.ceil(nx / ux) + 1
    ncy = 2 * math.ceil(ny / uy) + 1
    if mm.in_plane:
        uc_ox = mm.orientation[:uy, :ux, 0]
        uc_oy = mm.orientation[:uy, :ux, 1]
    else:
        uc_o = mm.occupation[:uy, :ux]
    uc_dx, uc_dy = np.meshgrid(mm.dx[:ux], mm.dy[:uy])
    tlm_dx = np.tile(uc_dx, (ncy, ncx))
    tlm_dy = np.tile(uc_dy, (ncy, ncx))
    index = -np.ones((uy, ux), dtype=int)
    Ks, Kps, Kpo = [], [], []
    for y in range(uy):
        for x in range(ux):
            sx = (ux - ((nx - 1) % ux) + x) % ux
            sy = (uy - ((ny - 1) % uy) + y) % uy
            sl = np.s_[sy:sy + 2 * ny - 1, sx:sx + 2 * nx - 1]
            dx, dy = tlm_dx[sl], tlm_dy[sl]
            xx, yy = np.cumsum(dx, axis=1), np.cumsum(dy, axis=0)
            rrx = xx - xx[0, nx - 1]
            rry = yy - yy[ny - 1, 0]
            with np.errstate(divide="ignore", invalid="ignore"):
                rr_sq = (rrx ** 2 + rry ** 2).astype(float)
                rr_sq[ny - 1, nx - 1] = np.inf
                rr_sq[np.where(rr_sq > cutoff_range ** 2)] = np.inf
                rr_inv = rr_sq ** -0.5
                rinv3 = rr_inv ** (-decay_exponent)
                rinv5 = rr_inv ** (-decay_exponent + 2)
                uxv = rrx * rr_inv
                uyv = rry * rr_inv
            if mm.in_plane:
                ox1, oy1 = uc_ox[y, x], uc_oy[y, x]
                if ox1 == oy1 == 0:
                    continue
                ox2 = np.tile(uc_ox, (ncy, ncx))[sl]
                oy2 = np.tile(uc_oy, (ncy, ncx))[sl]

                def pair(a1, b1, a2, b2):
                    k = a1 * a2 * (3 * uxv ** 2 - 1) + b1 * b2 * (3 * uyv ** 2 - 1) + 3 * (uxv * uyv) * (a1 * b2 + b1 * a2)
                    kc = a1 * a2 * (5 * uxv ** 2 - 1) + b1 * b2 * (5 * uyv ** 2 - 1) + 5 * (uxv * uyv) * (a1 * b2 + b1 * a2)
                    return -rinv3 * k - 3 * I / 2 * rinv5 * kc

                kernel = pair(ox1, oy1, ox2, oy2)
                kernel_ps = pair(-oy1, ox1, ox2, oy2)
                kernel_po = pair(ox1, oy1, -oy2, ox2)
            else:
                if uc_o[y, x] == 0:
                    continue
                o2 = np.tile(uc_o, (ncy, ncx)).astype(float)[sl]
                kernel = o2 * (rinv3 + 9 * I / 2 * rinv5)
                kernel_ps = np.zeros_like(kernel)
                kernel_po = np.zeros_like(kernel)
            if mm.PBC:
                for k in (kernel, kernel_ps, kernel_po):
                    _apply_pbc(k, nx, ny)
            index[y, x] = len(Ks)
            Ks.append(kernel * 1e-7)
            Kps.append(kernel_ps * 1e-7)
            Kpo.append(kernel_po * 1e-7)
    return np.asarray(Ks), np.asarray(Kps), np.asarray(Kpo), index


def _apply_pbc(k, nx, ny):
    """energies.py:264-273: add the 8 neighbouring images of the kernel."""
    kopy = k.copy()
    k[:, nx:] += kopy[:, :nx - 1]
    k[ny:, nx:] += kopy[:ny - 1, :nx - 1]
    k[ny:, :] += kopy[:ny - 1, :]
    k[ny:, :nx - 1] += kopy[:ny - 1, nx:]
    k[:, :nx - 1] += kopy[:, nx:]
    k[:ny - 1, :nx - 1] += kopy[ny:, nx:]
    k[:ny - 1, :] += kopy[ny:, :]
    k[:ny - 1, nx:] += kopy[ny:, :nx - 1]


# --------------------------------------------------------------------------------------------
# Magnets
# --------------------------------------------------------------------------------------------
DEFAULT_PARAMS = dict(
    SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF=20,
    REDUCED_KERNEL_SIZE=20,
    UPDATE_SCHEME="Neel",
    MULTISAMPLING_SCHEME="grid",
    ENERGY_BARRIER_METHOD="simple",
    METROPOLIS_TIME=False,
)


class OracleMagnets:
    """core.py:57-130 restated for the stepping path."""

    def __init__(self, geom, *, T=300, E_B=0.0, moment=None, Msat=800e3, V=2e-22, pattern,
                 energies, PBC=False, m_perp_factor=None, major_axis=0.0, minor_axis=None,
                 params=None, seed=0, rng="philox"):
        self.geom = geom
        self.nx, self.ny = geom.nx, geom.ny
        self.dx, self.dy = geom.dx, geom.dy
        self.x, self.y = geom.x, geom.y
        self.xx, self.yy = geom.xx, geom.yy
        self.ixx, self.iyy = geom.ixx, geom.iyy
        self.in_plane = geom.in_plane
        self.occupation = geom.occupation.copy()
        with np.errstate(divide="ignore"):
            self._occupation_inf = 1 / self.occupation
        self.nonzero = np.nonzero(self.occupation)
        self._nonzero_array = np.asarray(self.nonzero).reshape(2, -1)
        self.n = int(np.sum(self.occupation))
        self.orientation = geom.orientation
        self.ux, self.uy = geom.ux, geom.uy
        self.nn = geom.nn
        # core.py:361-375: PBC is refused (with a warning) unless whole unit cells tile the grid.
        self.PBC = bool(PBC) and self.nx % self.ux == 0 and self.ny % self.uy == 0
        self.params = dict(DEFAULT_PARAMS)
        if params:
            self.params.update(params)
        m = as_2D_array(pattern, (self.ny, self.nx))
        self.m = np.sign(m.astype(float)) * self.occupation
        self.shape = self.m.shape
        self.T = T
        self.E_B = as_2D_array(E_B, self.shape)
        with np.errstate(divide="ignore"):
            self._E_B_inv_over_four = 0.25 / self.E_B
        self.moment = as_2D_array(Msat * V if moment is None else moment, self.shape)
        self._momentSq = self.moment * self.moment
        self.m_perp_factor = float(self.in_plane if m_perp_factor is None else m_perp_factor)
        self.major_axis = major_axis
        self.minor_axis = major_axis if minor_axis is None else minor_axis
        self.t = 0.0
        self.switches = 0
        self.attempted_switches = 0
        self.seed = int(seed)
        self.step = 0
        self.rng_mode = rng
        if rng == "numpy":
            self._np_rng = np.random.default_rng(seed)
            self._np_legacy = np.random.RandomState(seed)
        self._calc_r_cache = {}
        self._zeros = np.zeros_like(self.m)
        self._energies = []
        for e in energies:
            e.initialize(self)
            self._energies.append(e)
        self.last_select = None

    # ---- fields ------------------------------------------------------------------------
    @property
    def T(self):
        return self._T

    @T.setter
    def T(self, value):
        self._T = as_2D_array(value, (self.ny, self.nx))
        self.kBT = kB * self._T
        self.beta = 1 / self.kBT

    @property
    def USE_PERP_ENERGY(self):
        return bool(self.m_perp_factor) and self.in_plane

    @property
    def MCsteps(self):
        return self.attempted_switches / self.n

    @property
    def E(self):
        return sum(e.E for e in self._energies)

    @property
    def E_perp(self):
        return sum(e.E_perp for e in self._energies) * self.m_perp_factor

    @property
    def E_tot(self):
        return float(sum(e.E_tot for e in self._energies))

    @property
    def m_avg_x(self):
        return float(np.mean(self.m * self.orientation[:, :, 0])) * self.nx * self.ny / self.n if self.in_plane else 0

    @property
    def m_avg_y(self):
        return float(np.mean(self.m * self.orientation[:, :, 1])) * self.nx * self.ny / self.n if self.in_plane else 0

    @property
    def m_avg(self):
        """core.py:419-421"""
        return (self.m_avg_x ** 2 + self.m_avg_y ** 2) ** 0.5 if self.in_plane else float(np.mean(self.m))

    def get_energy(self, name):
        name = name.lower().replace("energy", "")
        for e in self._energies:
            if type(e).__name__.lower() == name:
                return e
        return None

    def update_energy(self, index=None):
        """core.py:241-253"""
        for e in self._energies:
            if index is None:
                e.update()
            elif np.asarray(index[0]).size == 1:
                e.update_single(index)
            elif np.asarray(index[0]).size > 1:
                e.update_multiple(index)

    def switch_energy(self, idx=None):
        return sum(e.energy_switch(idx) for e in self._energies)

    def perp_energy(self, idx=None):
        if not self.USE_PERP_ENERGY:
            return self._zeros if idx is None else np.zeros(idx.shape[1])
        return self.m_perp_factor * sum(e.E_perp if idx is None else e.E_perp[idx[0], idx[1]] for e in self._energies)

    # ---- barriers --------------------------------------------------------------------
    def E_barrier(self, idx=None, min_only=False):
        """core.py:625-656"""
        with np.errstate(invalid="ignore", divide="ignore"):
            occ_inf = self._occupation_inf if idx is None else self._occupation_inf[idx[0], idx[1]]
            delta_E = self.switch_energy(idx) * occ_inf
            E = -delta_E / 2
            E_B = self.E_B if idx is None else self.E_B[idx[0], idx[1]]
            E_B_i4 = self._E_B_inv_over_four if idx is None else self._E_B_inv_over_four[idx[0], idx[1]]
            if not self.USE_PERP_ENERGY:
                if self.params["ENERGY_BARRIER_METHOD"] == "simple":
                    barrier = np.maximum(delta_E, E_B - E)
                else:
                    barrier = np.where(delta_E > 4 * E_B, delta_E,
                                       np.minimum(E_B * (delta_E * E_B_i4 + 1) ** 2, delta_E + 4 * E_B))
                return barrier if min_only else (barrier, barrier)
            E_perp = self.perp_energy(idx)
            b1 = np.maximum(delta_E, E_B + E_perp - E)
            b2 = np.maximum(delta_E, E_B - E_perp - E)
            return np.minimum(b1, b2) if min_only else (b1, b2)

    # ---- randomness hooks ------------------------------------------------------------
    def _exponential(self, scale):
        if self.rng_mode == "numpy":
            return self._np_rng.exponential(scale=scale, size=self.shape)
        u = philox.neel_uniforms(self.seed, self.step, self.nx * self.ny).reshape(self.shape)
        return scale * (-np.log(u))

    # ---- Neel --------------------------------------------------------------------------
    def _update_Neel(self, t_max=1, attempt_freq=1e10):
        """core.py:679-709; returns (y, x) or None. Also records the step's argmin diagnostics."""
        with np.errstate(invalid="ignore", divide="ignore", over="ignore"):
            if self.USE_PERP_ENERGY:
                b1, b2 = self.E_barrier()
                bmin = min(np.nanmin(b1), np.nanmin(b2))
                b1 = b1 - bmin
                b2 = b2 - bmin
                freq = np.exp(-b1 * self.beta) + np.exp(-b2 * self.beta)
                taus = self._exponential(1 / freq)
                attempt_freq /= 2
            else:
                b = self.E_barrier(min_only=True)
                bmin = np.nanmin(b)
                b = b - bmin
                taus = self._exponential(np.exp(b * self.beta))
            flat = int(np.nanargmin(taus))
            w = divmod(flat, self.nx)
            time = taus[w] * np.exp(bmin * self.beta[w]) / attempt_freq
            self.last_taus = taus
        self.attempted_switches += 1
        if time > t_max or np.isnan(time):
            self.t += t_max
            return None
        self.m[w] *= -1
        self.switches += 1
        self.t += float(time)
        self.update_energy(np.asarray(w).reshape(2, -1))
        return w

    # ---- Metropolis --------------------------------------------------------------------
    def calc_r(self, Q):
        """core.py:720-733 (cached per Q for the object's lifetime, as the reference's lru_cache)."""
        if Q not in self._calc_r_cache:
            with np.errstate(divide="ignore"):
                r = (8e-7 * self._momentSq / (Q * self.kBT)) ** (1 / 3)
            r = np.clip(r, 0, (self.geom.x_max - self.geom.x_min) + (self.geom.y_max - self.geom.y_min))
            self._calc_r_cache[Q] = float(np.max(r))
        return self._calc_r_cache[Q]

    def _draw_sel(self, attempt):
        if self.rng_mode == "numpy":
            return None
        return philox.sel_uniforms(self.seed, self.step, attempt)

    def _select_grid(self, r, attempt):
        """core.py:496-558 with the keyed draws; also returns each sample's supercell id."""
        Rx = math.ceil(r / np.min(self.dx)) - 1
        Ry = math.ceil(r / np.min(self.dy)) - 1
        Rx = self.ux * math.ceil(Rx / self.ux)
        Ry = self.uy * math.ceil(Ry / self.uy)
        Rx, Ry = min(Rx, self.nx), min(Ry, self.ny)
        Rx, Ry = max(Rx, 2), max(Ry, 2)
        snx = self.nx // Rx + 1
        sny = self.ny // Ry + 1
        sc_x = self.ixx[:sny:2, :snx:2].ravel()
        sc_y = self.iyy[:sny:2, :snx:2].ravel()
        ncols = self.ixx[:sny:2, :snx:2].shape[1]
        n = sc_x.size
        if self.rng_mode == "numpy":
            if self.PBC:
                ox, oy = self._np_legacy.randint(self.nx), self._np_legacy.randint(self.ny)
            else:
                ox, oy = self._np_legacy.randint(-Rx, Rx), self._np_legacy.randint(-Ry, Ry)
        else:
            u = philox.sel_uniforms(self.seed, self.step, attempt)
            if self.PBC:
                ox, oy = philox.randint(u[0], self.nx), philox.randint(u[1], self.ny)
            else:
                ox, oy = philox.randint(u[0], -Rx, Rx), philox.randint(u[1], -Ry, Ry)
        if Rx == self.nx:
            ox = 0
        if Ry == self.ny:
            oy = 0
        ddx, ddy = ox % self.ux, oy % self.uy
        occ_sc = self.occupation[ddy:ddy + Ry, ddx:ddx + Rx]
        nz = occ_sc.nonzero()
        cnt = nz[0].size
        if self.rng_mode == "numpy":
            pick = self._np_legacy.choice(cnt, n)
        else:
            cu = philox.choice_uniforms(self.seed, self.step, attempt, n)
            pick = np.floor(cu * cnt).astype(np.int64)
        idx_x = sc_x * Rx + nz[1][pick]
        idx_y = sc_y * Ry + nz[0][pick]
        ids = np.arange(n)
        if self.PBC:
            ok = (idx_x < max(Rx, self.nx - Rx)) & (idx_y < max(Ry, self.ny - Ry))
            idx_x, idx_y, ids = idx_x[ok], idx_y[ok], ids[ok]
            idx_x = (idx_x + ox) % self.nx
            idx_y = (idx_y + oy) % self.ny
        else:
            idx_x = idx_x + ox
            idx_y = idx_y + oy
            ok = (idx_x < self.nx) & (idx_y < self.ny) & (idx_x >= 0) & (idx_y >= 0)
            idx_x, idx_y, ids = idx_x[ok], idx_y[ok], ids[ok]
        self.last_select = dict(Rx=Rx, Ry=Ry, ox=ox, oy=oy, n=n, ncols=ncols)
        if idx_x.size != 0:
            return np.asarray([idx_y.reshape(-1), idx_x.reshape(-1)]), ids
        return self._select_single(attempt), np.zeros(1, dtype=np.int64)

    def _select_single(self, attempt):
        """core.py:489-494"""
        if self.rng_mode == "numpy":
            i = self._np_legacy.randint(self.n)
        else:
            i = philox.randint(philox.sel_uniforms(self.seed, self.step, attempt)[2], self.n)
        return self._nonzero_array[:, i].reshape(2, -1)

    def select(self, r):
        """core.py:456-487 ('grid' and 'single'); retries with a fresh attempt if nothing is valid."""
        attempt = 0
        while True:
            if self.params["MULTISAMPLING_SCHEME"] == "single":
                pos, ids = self._select_single(attempt), np.zeros(1, dtype=np.int64)
            else:
                pos, ids = self._select_grid(r, attempt)
            valid = np.where(self.occupation[pos[0], pos[1]])[0] if pos.size > 0 else np.zeros(0, dtype=int)
            if valid.size > 0:
                self.last_attempts = attempt + 1
                return pos[:, valid], ids[valid]
            attempt += 1

    def _update_Metropolis(self, idx=None, Q=0.05, r=0, attempt_freq=1e10):
        """core.py:659-677"""
        if idx is None:
            idx, ids = self.select(r=self.calc_r(Q) if r == 0 else r)
        else:
            idx = np.asarray(idx)
            ids = np.arange(idx.shape[1])
        self.attempted_switches += idx.shape[1]
        beta = self.beta[idx[0], idx[1]]
        with np.errstate(divide="ignore", over="ignore"):
            expo = np.clip(np.exp(-self.switch_energy(idx) * beta), 1e-10, 1e10)
            if self.params["METROPOLIS_TIME"]:
                if self.rng_mode == "numpy":
                    u = self._np_rng.random(size=idx.shape[1])
                else:
                    u = philox.mtime_uniforms(self.seed, self.step, idx[0] * self.nx + idx[1])
                times = -np.log(u) / attempt_freq / self.n * np.exp(self.E_barrier(idx, min_only=True) * beta)
                self.t += np.sum(times)
        if self.rng_mode == "numpy":
            ua = self._np_rng.random(size=expo.shape)
        else:
            ua = philox.accept_uniforms(self.seed, self.step, idx[0] * self.nx + idx[1])
        self.last_samples = (idx.copy(), ids.copy(), ua, expo)
        idx_switch = idx[:, np.where(ua < expo)[0]]
        if idx_switch.shape[1] > 0:
            self.m[idx_switch[0], idx_switch[1]] *= -1
            self.switches += idx_switch.shape[1]
            self.update_energy(index=idx_switch)
        return idx_switch

    # ---- dispatch ----------------------------------------------------------------------
    def update(self, *args, **kwargs):
        """core.py:604-622 (one step; the step counter keys the random streams)."""
        if np.any(self.T == 0):
            return None
        if self.params["UPDATE_SCHEME"] in ("Neel", "Néel"):
            out = self._update_Neel(*args, **kwargs)
        elif self.params["UPDATE_SCHEME"] == "Wolff":
            out = self._update_Wolff(*args, **kwargs)
        else:
            out = self._update_Metropolis(*args, **kwargs)
        self.step += 1
        return out

    def _select_cluster_exchange(self):
        """core.py:588-602. Philox mode: seed = _select_single's draw, bonds on the uniform
        (cell, step, tag 7 | iteration << 8); NumPy mode: the reference's own draws."""
        ex = self.get_energy("exchange")
        neighbors = ex.local_interaction
        seed = tuple(self._select_single(0).flat)
        cluster = np.zeros_like(self.m)
        cluster[seed] = 1
        checked = np.copy(cluster)
        checked[np.where(self.m != self.m[seed])] = 1
        bnd = "wrap" if self.PBC else "fill"
        it = 0
        while np.any(frontier := (1 - checked) * signal.convolve2d(cluster, neighbors, mode="same", boundary=bnd)):
            checked[frontier != 0] = 1
            bond_prob = 1 - np.exp(-2 * ex.J / self.kBT * frontier)
            if self.rng_mode == "numpy":
                u = self._np_rng.random(size=cluster.shape)
            else:
                u = philox.uniform53(self.seed, self.step, philox.TAG_WOLFF | (it << 8),
                                     np.arange(self.nx * self.ny)).reshape(cluster.shape)
            frontier[u > bond_prob] = 0
            cluster[frontier != 0] = 1
            it += 1
        return np.asarray(np.where(cluster == 1))

    def _update_Wolff(self):
        """core.py:711-718"""
        cluster = self._select_cluster_exchange()
        self.m[cluster[0], cluster[1]] *= -1
        self.switches += cluster.shape[1]
        self.attempted_switches += cluster.shape[1]
        self.update_energy(index=cluster)
        return cluster

    def progress(self, t_max=1, MCsteps_max=4.0, **kwargs):
        """core.py:736-760"""
        t0, mc0 = self.t, self.MCsteps
        while lower_than(dt := self.t - t0, t_max) and lower_than(self.MCsteps - mc0, MCsteps_max):
            if self.params["UPDATE_SCHEME"] in ("Neel", "Néel"):
                self.update(t_max=(t_max - dt), **kwargs)
            else:
                self.update(**kwargs)
        return self.t - t0, self.MCsteps - mc0

    # ---- zero-temperature relaxation ----------------------------------------------------
    def minimize(self, ignore_barrier=True, _frac=0.3):
        """core.py:763-783 (ignore_barrier=False ranks by the smaller E_barrier)."""
        visited = np.zeros_like(self.m)
        while True:
            barrier = self.switch_energy() if ignore_barrier else np.where(self.m != 0, self.E_barrier(min_only=True),
                                                                           np.inf)
            nobarrier = np.where(barrier < -self.kBT)
            if (n := nobarrier[0].size) != 0:
                rand = np.random.choice(n, size=math.ceil(n * _frac), replace=False)
                idx = (nobarrier[0][rand], nobarrier[1][rand])
            else:
                idx = divmod(int(np.argmin(barrier)), self.nx)
                if barrier[idx] >= 0:
                    break
                visited[idx] += 1
                if visited[idx] >= 3:
                    break
            idx = np.asarray(idx).reshape(2, -1)
            self.m[idx[0], idx[1]] *= -1
            self.switches += idx.shape[1]
            self.attempted_switches += idx.shape[1]
            self.update_energy(index=idx)

    def _minimize_all(self, ignore_barrier=True, _simultaneous=False, order=None):
        """core.py:785-806. `order`: the subset permutation (default: numpy.random.permutation,
        as the reference)."""
        if _simultaneous:
            step_x = step_y = 1
        else:
            step_x, step_y = self.ux * 2, self.uy * 2
        if order is None:
            order = np.random.permutation(step_x * step_y)
        for i in order:
            y, x = divmod(int(i), step_x)
            if not self.occupation[y, x]:
                continue
            sub = np.asarray(np.meshgrid(np.arange(y, self.ny, step_y), np.arange(x, self.nx, step_x))).reshape(2, -1)
            if ignore_barrier:
                idx = sub[:, np.where(self.switch_energy(sub) < 0)[0]]
            else:
                idx = sub[:, self.switch_energy(sub) / 2 + self.E_B[sub[0], sub[1]] < 0]
            if idx.size > 0:
                self.m[idx[0], idx[1]] = -self.m[idx[0], idx[1]]
                self.switches += idx.shape[1]
                self.attempted_switches += idx.shape[1]
                self.update_energy(index=idx)

    def relax(self, **kwargs):
        """core.py:808-845 (without the plotting)."""
        n, n_max = 0, int(math.sqrt(self.nx ** 2 + self.ny ** 2))
        prev = [self.m.copy(), self._zeros, self._zeros]
        switches = self.switches
        while not np.allclose(prev[2], prev[0]) and n < n_max:
            n += 1
            self._minimize_all(**kwargs)
            if switches == self.switches:
                break
            switches = self.switches
            prev[2] = prev[1].copy()
            prev[1] = prev[0].copy()
            prev[0] = self.m.copy()
        return n


def make_magnets(lattice, a, n=None, nx=None, ny=None, angle=0.0, PBC=False, **kw):
    """Convenience: geometry + OracleMagnets."""
    g = lattices.make(lattice, a, n=n, nx=nx, ny=ny, angle=angle, PBC=PBC)
    return OracleMagnets(g, PBC=PBC, **kw)
