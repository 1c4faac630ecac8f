"""Host helpers of the stepping loop (no GPU)."""
import itertools
import math

import numpy as np
import pytest


def test_lower_than_is_numpys_isclose_rule():
    """utils.lower_than (hotspice/utils.py:90-92) evaluates np.isclose(x, y, atol=0, rtol) in scalar
    form; it must agree with NumPy on every pair, including signed zeros, infinities, NaN and
    values within and just outside the tolerance."""
    from hotspice_amd.utils import lower_than
    vals = [0.0, -0.0, 1.0, 1.0 + 1e-11, 1.0 + 1e-9, 1.0 - 1e-11, -1.0, 5e-324, 1e308, -1e308,
            math.inf, -math.inf, math.nan, 2.0, 1e-20, 3, 7]
    rng = np.random.default_rng(0)
    vals += list(rng.normal(size=30)) + list(1 + rng.normal(size=30) * 1e-10)
    for x, y in itertools.product(vals, vals):
        ref = bool(x < y and not np.isclose(x, y, atol=0, rtol=1e-10))
        assert lower_than(x, y) == ref, (x, y)
    assert lower_than(1.0, 1.0 + 1e-3, rtol=1e-2) is False and lower_than(1.0, 1.1, rtol=1e-2) is True


def test_region_bins_follow_the_reference_linspace():
    """Readout bins: the reference bins cell i into int(linspace(0, nb*(1-1e-15), n)[i])
    (io.py:347-348); the device readout sums contiguous ranges, whose bounds must give exactly
    those bin indices."""
    from hotspice_amd.io import _bins
    for nb, nc in [(1, 7), (3, 10), (4, 4), (5, 64), (7, 100), (16, 1023), (2, 3)]:
        idx, starts = _bins(nb, nc)
        ref = np.linspace(0, nb * (1 - 1e-15), nc).astype(int)
        assert np.array_equal(idx, ref)
        assert starts[0] == 0 and starts[-1] == nc and np.all(np.diff(starts) >= 0)
        for b in range(nb):
            assert np.all(ref[starts[b]:starts[b + 1]] == b)


def test_inputter_transforms_and_bits():
    """OOPSquareChessStepsInputter's value-to-magnitude map (io.py:489-500) at its knots, and the
    little-endian bits of IntegerDatastream.as_bits (io.py:56-62)."""
    from hotspice_amd import io
    inp = io.OOPSquareChessStepsInputter(io.RandomScalarDatastream(), magnitude=2., magnitude_range=.5,
                                         transition_range=.4)
    s = 0.5 - 0.4 / 2
    assert inp.transform(0.) == -2.
    assert abs(inp.transform(s) - (-2. + .5)) < 1e-15
    assert inp.transform(1.) == 2.
    assert inp.transform(0.5) == 0.
    assert abs(inp.transform(1 - s) - (2. - .5)) < 1e-12
    binp = io.OOPSquareChessStepsInputter(io.RandomBinaryDatastream(), magnitude=3.)
    assert binp.transform(1) == 3. and binp.transform(0) == -3.
    ds = io.RandomIntegerDatastream(num_bits=5)
    assert ds.as_bits(6).tolist() == [0, 1, 1, 0, 0]
    assert ds.as_bits(6, endianness='big').tolist() == [0, 0, 1, 1, 0]
    ds.reset_rng()
    a = ds.get_next(8)
    ds.reset_rng()
    assert np.array_equal(a, ds.get_next(8)) and a.max() < 32
    c = io.ClockingFieldInputter(io.RandomBinaryDatastream(), angle=0.3, spread=0.1)
    assert c.bit_to_angles(0) == (0.3 + 0.1, 0.3 - 0.1) and c.bit_to_angles(1) == (0.3 + 0.1 + np.pi, 0.3 - 0.1 + np.pi)


def test_unit_helpers():
    from hotspice_amd import utils
    assert utils.eV_to_J(1.0) == 1.602176634e-19
    assert utils.J_to_eV(utils.eV_to_J(2.5)) == pytest.approx(2.5, rel=1e-15)
    a = np.arange(4.0)
    assert np.array_equal(utils.asnumpy(a), a)
