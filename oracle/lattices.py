"""Lattice geometry and unit-cell detection for the oracle (NumPy, FP64).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): imported by tests, smoke() and the
cpu_baseline leg of bench.py, never by the product.

Restates the geometry of the lattices the stepping configurations use:
  OOP_Square        hotspice/ASI.py:48-78
  IP_Square_Closed  hotspice/ASI.py:266-301 (IP_Square alias, ASI.py:317)
  IP_Pinwheel       hotspice/ASI.py:304-314, 357 (IP_Square rotated by -pi/4)
  IP_Kagome         hotspice/ASI.py:371-414
  IP_Ising          hotspice/ASI.py:230-263
and the unit-cell search of Magnets._get_unitcell (hotspice/core.py:146-161) with
check_repetition (hotspice/utils.py:42-66).
"""
import math

import numpy as np


class Geometry:
    """Plain container: grid, spacings, occupation, orientation, unit cell, NN mask."""

    def __init__(self, name, nx, ny, dx, dy, in_plane, occupation, angles, nn):
        self.name = name
        self.nx, self.ny = int(nx), int(ny)
        self.in_plane = bool(in_plane)
        # Magnets.dx/dy setters: tile the given spacing(s) to nx / ny entries (core.py:382-411)
        dxv = np.asarray(dx, dtype=float).reshape(-1)
        dyv = np.asarray(dy, dtype=float).reshape(-1)
        self.dx = np.tile(dxv, math.ceil(self.nx / dxv.size))[: self.nx]
        self.dy = np.tile(dyv, math.ceil(self.ny / dyv.size))[: self.ny]
        self.x = np.zeros(self.nx)
        self.x[1:] = np.cumsum(self.dx)[:-1]
        self.y = np.zeros(self.ny)
        self.y[1:] = np.cumsum(self.dy)[:-1]
        self.xx, self.yy = np.meshgrid(self.x, self.y)
        self.ixx, self.iyy = np.meshgrid(np.arange(self.nx), np.arange(self.ny))
        self.occupation = occupation.astype(bool).astype(int)
        self.nn = np.asarray(nn)
        if self.in_plane:
            ang = angles * self.occupation
            self.angles = ang
            self.orientation = np.zeros((self.ny, self.nx, 2))
            self.orientation[:, :, 0] = np.cos(ang) * self.occupation
            self.orientation[:, :, 1] = np.sin(ang) * self.occupation
        else:
            self.angles = None
            self.orientation = None
        self.ux, self.uy = unitcell(self)

    @property
    def x_min(self):
        return float(self.x[0])

    @property
    def x_max(self):
        return float(self.x[-1])

    @property
    def y_min(self):
        return float(self.y[0])

    @property
    def y_max(self):
        return float(self.y[-1])


def check_repetition(arr, nx, ny):
    """True if `arr` is periodic with period nx along axis 1 and ny along axis 0 (utils.py:42-66)."""
    extra = [1] * (arr.ndim - 2)
    max_y, max_x = arr.shape[:2]
    i, cur = 0, arr[:ny, :nx, ...]
    end_y, end_x = cur.shape[:2]
    while end_y < arr.shape[0] or end_x < arr.shape[1]:
        if i % 2:
            cur = np.tile(cur, (1, 2, *extra))
            start_x = cur.shape[1] // 2
            end_y, end_x = cur.shape[:2]
            if not np.allclose(cur[:max_y, start_x:max_x, ...], arr[:end_y, start_x:end_x, ...]):
                return False
        else:
            cur = np.tile(cur, (2, 1, *extra))
            start_y = cur.shape[0] // 2
            end_y, end_x = cur.shape[:2]
            if not np.allclose(cur[start_y:max_y, :max_x, ...], arr[start_y:end_y, :end_x, ...]):
                return False
        i += 1
    return True


def unitcell(g, max_cell=100):
    """Triangle search over (ux, uy) as in Magnets._get_unitcell (core.py:146-161)."""
    dx, dy = g.dx.reshape(1, -1), g.dy.reshape(-1, 1)
    arr = g.orientation if g.in_plane else g.occupation
    for n in range(1, min(g.nx + g.ny + 1, max_cell + 1)):
        for uy in range(max(0, n - g.nx) + 1, min(n, g.ny) + 1):
            ux = n - uy + 1
            if check_repetition(arr, ux, uy):
                if check_repetition(dx, ux, 1) and check_repetition(dy, 1, uy):
                    return ux, uy
    return g.nx, g.ny


def make(name, a, n=None, nx=None, ny=None, angle=0.0, PBC=False):
    """Geometry of a named lattice with the reference's default spacings."""
    nx = n if nx is None else nx
    key = name.lower()
    if key == "oop_square":
        ny = n if ny is None else ny
        occ = np.ones((ny, nx), dtype=int)
        return Geometry(name, nx, ny, a, a, False, occ, None, [[0, 1, 0], [1, 0, 1], [0, 1, 0]])
    if key in ("ip_square", "ip_square_closed", "ip_pinwheel", "ip_pinwheel_diamond"):
        ny = n if ny is None else ny
        ixx, iyy = _grid(nx, ny)
        occ = ((ixx + iyy) % 2 == 1).astype(int)
        ang = np.zeros((ny, nx))
        ang[iyy % 2 == 1] = math.pi / 2
        extra = angle - (math.pi / 4 if "pinwheel" in key else 0.0)
        return Geometry(name, nx, ny, a / 2, a / 2, True, occ, ang + extra, [[1, 0, 1], [0, 0, 0], [1, 0, 1]])
    if key == "ip_ising":
        ny = n if ny is None else ny
        occ = np.ones((ny, nx), dtype=int)
        return Geometry(name, nx, ny, a, a, True, occ, np.zeros((ny, nx)) + angle, [[0, 1, 0], [1, 0, 1], [0, 1, 0]])
    if key == "ip_kagome":
        if ny is None:
            ny = int(nx / math.sqrt(3)) // 4 * 4
            if not PBC:
                ny -= 1
        ixx, iyy = _grid(nx, ny)
        occ = np.zeros((ny, nx), dtype=int)
        occ[(ixx + iyy) % 4 == 1] = 1
        occ[(ixx - iyy) % 4 == 3] = 1
        ang = np.ones((ny, nx)) * math.pi / 2
        ang[((ixx - iyy) % 4 == 1) & (ixx % 2 == 1)] = -math.pi / 6
        ang[((ixx + iyy) % 4 == 3) & (ixx % 2 == 1)] = math.pi / 6
        dx = a / 4
        return Geometry(name, nx, ny, dx, math.sqrt(3) * dx, True, occ, ang + angle,
                        [[0, 1, 0, 1, 0], [1, 0, 0, 0, 1], [0, 1, 0, 1, 0]])
    raise ValueError(f"oracle: unknown lattice {name}")


def _grid(nx, ny):
    return np.meshgrid(np.arange(nx), np.arange(ny))
