"""Philox4x32-10 counter-based generator and the uniform streams of the stepping path (NumPy).

TEST INFRASTRUCTURE ONLY. This module is part of the CPU oracle: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import it,
and only as the checker. The product (``hotspice_amd``) never imports it.

The generator follows Salmon, Moraes, Dror & Shaw, "Parallel random numbers: as easy as
1, 2, 3" (SC'11): multipliers 0xD2511F53 / 0xCD9E8D57, Weyl key increments
0x9E3779B9 / 0xBB67AE85, ten rounds, key bumped between rounds. It is pinned by the
published known-answer vectors of that paper's reference implementation (tests/test_oracle_philox.py).

Stream contract (shared bit-for-bit with the HIP kernels in hotspice_amd/csrc):
  key     = (seed & 0xffffffff, seed >> 32)
  counter = (index, step & 0xffffffff, step >> 32, tag)
Each Philox call yields four 32-bit words; a 53-bit uniform takes two of them, so one
call serves two consecutive indices (words 0,1 for even, 2,3 for odd index).
  u53 in [0, 1):  ((w0 >> 5) * 2**26 + (w1 >> 6)) * 2**-53
  u53 in (0, 1]:  the same plus 2**-53
The reference draws from NumPy's PCG64/MT19937 (hotspice/core.py:83, 493, 527-538, 671,
698); those streams are not reproducible (unseeded), so the build replaces them with these
keyed streams and replays them into the reference for parity.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = 0x9E3779B9
W1 = 0xBB67AE85
MASK32 = np.uint64(0xFFFFFFFF)

# Stream tags (counter word 3). Attempt numbers of the Metropolis selection go in bits 8..23.
TAG_NEEL = 1      # Neel: one exponential per grid cell per step (core.py:692, 698)
TAG_SEL = 2       # grid selection: supergrid offsets and single-magnet fallback (core.py:493, 527-529)
TAG_CHOICE = 3    # grid selection: occupied cell inside each supercell (core.py:538)
TAG_ACCEPT = 4    # Metropolis acceptance, one per sample, indexed by its grid cell (core.py:671)
TAG_MTIME = 5     # METROPOLIS_TIME exponentials, indexed by the sample's grid cell (core.py:668)
TAG_WOLFF = 7     # Wolff bonds: one per grid cell per cluster iteration, tag 7 | iteration << 8 (core.py:599)

TWO_M53 = 2.0 ** -53


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32-R. Inputs broadcast; returns four uint64 arrays holding 32-bit words."""
    c0 = np.asarray(c0, dtype=np.uint64) & MASK32
    c1 = np.asarray(c1, dtype=np.uint64) & MASK32
    c2 = np.asarray(c2, dtype=np.uint64) & MASK32
    c3 = np.asarray(c3, dtype=np.uint64) & MASK32
    c0, c1, c2, c3 = np.broadcast_arrays(c0, c1, c2, c3)
    k0 = np.uint64(int(k0) & 0xFFFFFFFF)
    k1 = np.uint64(int(k1) & 0xFFFFFFFF)
    for r in range(rounds):
        if r:
            k0 = np.uint64((int(k0) + W0) & 0xFFFFFFFF)
            k1 = np.uint64((int(k1) + W1) & 0xFFFFFFFF)
        p0 = M0 * c0
        p1 = M1 * c2
        c0, c1, c2, c3 = (p1 >> np.uint64(32)) ^ c1 ^ k0, p1 & MASK32, (p0 >> np.uint64(32)) ^ c3 ^ k1, p0 & MASK32
    return c0, c1, c2, c3


def _key(seed):
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    return seed & 0xFFFFFFFF, seed >> 32


def _u53_words(w0, w1):
    return ((w0 >> np.uint64(5)) * np.uint64(1 << 26) + (w1 >> np.uint64(6))).astype(np.float64)


def uniform53(seed, step, tag, index, open_zero=False):
    """53-bit uniforms for integer array ``index`` under (seed, step, tag).

    ``open_zero=False`` gives [0,1); ``True`` gives (0,1].
    """
    index = np.asarray(index, dtype=np.int64)
    k0, k1 = _key(seed)
    step = int(step)
    w = philox4x32(index >> 1, step & 0xFFFFFFFF, step >> 32, int(tag), k0, k1)
    odd = (index & 1).astype(bool)
    a = np.where(odd, w[2], w[0])
    b = np.where(odd, w[3], w[1])
    x = _u53_words(a, b)
    if open_zero:
        x = x + 1.0
    return x * TWO_M53


def neel_uniforms(seed, step, ncells):
    """(0,1] uniform per grid cell (flat row-major index) for one Neel step."""
    return uniform53(seed, step, TAG_NEEL, np.arange(ncells), open_zero=True)


def sel_uniforms(seed, step, attempt):
    """Three [0,1) uniforms of one grid-selection attempt: offset_x, offset_y, single-magnet pick."""
    u = uniform53(seed, step, TAG_SEL, np.array([4 * attempt, 4 * attempt + 1, 4 * attempt + 2]))
    return float(u[0]), float(u[1]), float(u[2])


def choice_uniforms(seed, step, attempt, n):
    """[0,1) uniform per supercell j of one selection attempt."""
    return uniform53(seed, step, TAG_CHOICE | (int(attempt) << 8), np.arange(n))


def accept_uniforms(seed, step, cells):
    """[0,1) acceptance uniform per sample, keyed by the sample's flat grid cell y*nx+x."""
    return uniform53(seed, step, TAG_ACCEPT, np.asarray(cells))


def mtime_uniforms(seed, step, cells):
    """(0,1] uniform per sample for METROPOLIS_TIME, keyed by the sample's flat grid cell."""
    return uniform53(seed, step, TAG_MTIME, np.asarray(cells), open_zero=True)


def randint(u, lo, hi=None):
    """Integer in [lo, hi) from a [0,1) uniform: lo + floor(u*(hi-lo)) (NumPy-style signature)."""
    if hi is None:
        lo, hi = 0, lo
    return int(lo + np.floor(u * (hi - lo)))
