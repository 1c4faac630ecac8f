"""Lattices of the stepping configurations (hotspice/ASI.py).

The geometry of each lattice -- which grid cells hold a magnet, their easy-axis angles, the
default spacings, the nearest-neighbour mask used by ExchangeEnergy and the named initial
patterns -- is a data format fixed by the reference, so it is reproduced as-is:
  OOP_Square   ASI.py:48-100      IP_Ising        ASI.py:230-263
  IP_Square    ASI.py:266-301,317 IP_Pinwheel     ASI.py:304-314,357
  IP_Kagome    ASI.py:371-414
"""
import math

import numpy as np

from .core import Magnets


def _signal_conv_same(a, k, wrap):
    """2D convolution, output the size of `a` (fill with zeros, or wrap), for small masks."""
    a = np.asarray(a, dtype=float)
    k = np.asarray(k, dtype=float)
    hy, hx = k.shape[0] // 2, k.shape[1] // 2
    out = np.zeros_like(a)
    ny, nx = a.shape
    for qy in range(k.shape[0]):
        for qx in range(k.shape[1]):
            w = k[qy, qx]
            if w == 0:
                continue
            sy, sx = hy - qy, hx - qx  # out[i] += k[q] * a[i + h - q]
            if wrap:
                out += w * np.roll(a, (-sy, -sx), axis=(0, 1))
            else:
                src = np.zeros_like(a)
                ys, ye = max(0, sy), min(ny, ny + sy)
                xs, xe = max(0, sx), min(nx, nx + sx)
                src[ys - sy:ye - sy, xs - sx:xe - sx] = a[ys:ye, xs:xe]
                out += w * src
    return out


class OOP_ASI(Magnets):
    """Out-of-plane ASI (ASI.py:15-36)."""

    def _get_angles(self):
        return np.nan * np.zeros_like(self.ixx)

    def correlation_NN(self, NN=None):
        """ASI.py:28-36: <S_i S_j> over nearest neighbours."""
        if NN is None:
            NN = self._get_nearest_neighbors()
        total = _signal_conv_same(self.occupation, NN, self.PBC) * self.occupation
        m = self.m
        nsum = _signal_conv_same(m, NN, self.PBC)
        valid = np.where(total != 0)
        return float(np.mean(m[valid] * nsum[valid] / total[valid]))


class IP_ASI(Magnets):
    """In-plane ASI (ASI.py:38-45)."""


class OOP_Square(OOP_ASI):
    def __init__(self, a: float, n: int = None, *, nx: int = None, ny: int = None, **kwargs):
        """ASI.py:48-56"""
        self.a = a
        nx = n if nx is None else nx
        ny = n if ny is None else ny
        if nx is None or ny is None:
            raise AttributeError("Must specify <n> if either <nx> or <ny> are not specified.")
        dx, dy = kwargs.pop('dx', a), kwargs.pop('dy', a)
        super().__init__(nx, ny, dx, dy, in_plane=False, **kwargs)

    def _set_m(self, pattern: str):
        if str(pattern).strip().lower() == 'afm':
            self._m_pattern = ((self.ixx - self.iyy) % 2) * 2 - 1
        else:
            super()._set_m(pattern)

    def _get_occupation(self):
        return np.ones_like(self.ixx)

    def _get_nearest_neighbors(self):
        return np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]])

    def _get_groundstate(self):
        return 'afm'

    def _get_appropriate_avg(self):
        return ['point', 'square', 'crossfour']

    def correlation_NN(self, N=1):
        """ASI.py:93-100"""
        masks = {1: None, 2: np.array([[1, 0, 1], [0, 0, 0], [1, 0, 1]]),
                 3: np.array([[0, 0, 1, 0, 0], [0, 0, 0, 0, 0], [1, 0, 0, 0, 1], [0, 0, 0, 0, 0], [0, 0, 1, 0, 0]])}
        if N not in masks:
            raise ValueError("Can only calculate NN correlation up to third-nearest neighbor")
        return super().correlation_NN(NN=masks[N])


class IP_Ising(IP_ASI):
    def __init__(self, a: float, n: int = None, *, nx: int = None, ny: int = None, **kwargs):
        """ASI.py:230-238"""
        self.a = a
        nx = n if nx is None else nx
        ny = n if ny is None else ny
        if nx is None or ny is None:
            raise AttributeError("Must specify <n> if either <nx> or <ny> are not specified.")
        dx, dy = kwargs.pop('dx', a), kwargs.pop('dy', a)
        super().__init__(nx, ny, dx, dy, in_plane=True, **kwargs)

    def _set_m(self, pattern: str):
        if str(pattern).strip().lower() == 'afm':
            self._m_pattern = (self.iyy % 2) * 2 - 1
        else:
            super()._set_m(pattern)

    def _get_angles(self):
        return np.zeros_like(self.xx)

    def _get_occupation(self):
        return np.ones_like(self.ixx)

    def _get_nearest_neighbors(self):
        return np.array([[0, 1, 0], [1, 0, 1], [0, 1, 0]])

    def _get_groundstate(self):
        return 'uniform'


class IP_Square_Closed(IP_ASI):
    def __init__(self, a: float, n: int = None, *, nx: int = None, ny: int = None, **kwargs):
        """ASI.py:266-274: magnets on the edges of squares of side a (cells a/2 apart)."""
        self.a = a
        nx = n if nx is None else nx
        ny = n if ny is None else ny
        if nx is None or ny is None:
            raise AttributeError("Must specify <n> if either <nx> or <ny> are not specified.")
        dx, dy = kwargs.pop('dx', a / 2), kwargs.pop('dy', a / 2)
        super().__init__(nx, ny, dx, dy, in_plane=True, **kwargs)

    def _set_m(self, pattern: str):
        if str(pattern).strip().lower() == 'afm':
            self._m_pattern = ((self.ixx - self.iyy) // 2 % 2) * 2 - 1
        else:
            super()._set_m(pattern)

    def _get_angles(self):
        angles = np.zeros_like(self.xx)
        angles[self.iyy % 2 == 1] = math.pi / 2
        return angles

    def _get_occupation(self):
        return (self.ixx + self.iyy) % 2 == 1

    def _get_nearest_neighbors(self):
        return np.array([[1, 0, 1], [0, 0, 0], [1, 0, 1]])

    def _get_groundstate(self):
        return 'afm'

    def vertex_counts(self):
        """Fractions (N1, N2, N3, N4) of the four square-ice vertex types, counted as in
        examples/IP_SquareVertices.py:19-35: at every vertex (even x and y), the in-pointing count
        is the wrapped correlation of m with [[0,1,0],[1,0,-1],[0,-1,0]] (|.| = 2: type III,
        4: type IV) and the 'flower' sum the correlation with [[0,1,0],[-1,0,1],[0,-1,0]]
        (0 in-pointing and |flower| = 4: type I, otherwise type II)."""
        m = self.m

        def corr_wrap(mask):  # convolve2d(m, mask[::-1, ::-1], 'same', boundary='wrap')
            out = np.zeros_like(m)
            for i in range(3):
                for j in range(3):
                    if mask[i][j]:
                        out += mask[i][j] * np.roll(m, (1 - i, 1 - j), axis=(0, 1))
            return out

        vertices = (self.ixx % 2 == 0) & (self.iyy % 2 == 0)
        n_in = corr_wrap([[0, 1, 0], [1, 0, -1], [0, -1, 0]])[vertices]
        flower = corr_wrap([[0, 1, 0], [-1, 0, 1], [0, -1, 0]])[vertices]
        n3 = np.sum(np.abs(n_in) == 2)
        n4 = np.sum(np.abs(n_in) == 4)
        n1 = np.sum((np.abs(n_in) == 0) & (np.abs(flower) == 4))
        n2 = np.sum((np.abs(n_in) == 0) & (np.abs(flower) != 4))
        total = n1 + n2 + n3 + n4
        return np.array([n1, n2, n3, n4], dtype=float) / total if total else np.zeros(4)


class IP_Pinwheel_Diamond(IP_Square_Closed):
    def __init__(self, a: float, n: int = None, *, nx: int = None, ny: int = None, **kwargs):
        """ASI.py:304-308: IP_Square with every magnet rotated by -45 degrees."""
        kwargs['angle'] = kwargs.get('angle', 0) - math.pi / 4
        super().__init__(a, n=n, nx=nx, ny=ny, **kwargs)

    def _get_groundstate(self):
        return 'uniform' if self.PBC else 'vortex'


class IP_Square(IP_Square_Closed):
    pass


class IP_Pinwheel(IP_Pinwheel_Diamond):
    pass


class IP_Kagome(IP_ASI):
    def __init__(self, a: float, n: int = None, *, nx: int = None, ny: int = None, **kwargs):
        """ASI.py:371-382: a = distance between opposing sides of a hexagon."""
        self.a = a
        nx = n if nx is None else nx
        if ny is None and nx is not None:
            ny = int(nx / math.sqrt(3)) // 4 * 4
            if not kwargs.get('PBC', False):
                ny -= 1
        if nx is None or ny is None:
            raise AttributeError("Must specify <n> if either <nx> or <ny> are not specified.")
        dx = kwargs.pop('dx', a / 4)
        dy = kwargs.pop('dy', math.sqrt(3) * dx)
        super().__init__(nx, ny, dx, dy, in_plane=True, **kwargs)

    def _set_m(self, pattern: str):
        if str(pattern).strip().lower() == 'afm':
            m = np.ones_like(self.ixx)
            m[(self.ixx + self.iyy) % 4 == 3] *= -1
            self._m_pattern = m
        else:
            super()._set_m(pattern)

    def _get_angles(self):
        angles = np.ones_like(self.xx) * math.pi / 2
        angles[((self.ixx - self.iyy) % 4 == 1) & (self.ixx % 2 == 1)] = -math.pi / 6
        angles[((self.ixx + self.iyy) % 4 == 3) & (self.ixx % 2 == 1)] = math.pi / 6
        return angles

    def _get_occupation(self):
        occ = np.zeros_like(self.ixx)
        occ[(self.ixx + self.iyy) % 4 == 1] = 1
        occ[(self.ixx - self.iyy) % 4 == 3] = 1
        return occ

    def _get_nearest_neighbors(self):
        return np.array([[0, 1, 0, 1, 0], [1, 0, 0, 0, 1], [0, 1, 0, 1, 0]])

    def _get_groundstate(self):
        return 'uniform'
