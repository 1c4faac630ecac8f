"""Host helpers of the stepping loop (no GPU)."""
import itertools
import math

import numpy as np


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
