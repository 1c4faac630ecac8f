"""Host-side helpers the stepping path's API needs (hotspice/utils.py:42-92, 207-228)."""
import math

import numpy as np


def as_2D_array(value, shape) -> np.ndarray:
    """utils.py:208-228: scalar -> constant array, array -> reshaped (size must match)."""
    is_scalar = True
    if isinstance(value, list):
        try:
            value = np.asarray(value, dtype=float)
        except ValueError:
            raise ValueError("List of lists could not be converted to rectangular float64 array.")
    if isinstance(value, np.ndarray):
        if value.size != 1:
            is_scalar = False
    elif hasattr(value, "cpu") and hasattr(value, "numel"):  # torch tensor
        value = value.detach().cpu().numpy()
        if value.size != 1:
            is_scalar = False
    if is_scalar:
        return np.ones(shape, dtype=float) * float(np.asarray(value).reshape(-1)[0] if np.ndim(value) else value)
    if value.size != math.prod(shape):
        raise ValueError(f"Incorrect shape: passed array of shape {value.shape} is not of desired shape {shape}.")
    return np.asarray(value).reshape(shape)


def lower_than(x, y, rtol=1e-10):
    """utils.py:90-92: False if x is within rtol of y or greater than y. Scalar form of
    np.isclose(x, y, atol=0, rtol=rtol): |x - y| <= rtol*|y| for finite y, x == y otherwise (the
    host loops call this once per instance per launch, where np.isclose costs microseconds)."""
    if not x < y:
        return False
    x, y = float(x), float(y)
    if math.isinf(y):
        return True  # x < y, so x != y
    return not abs(x - y) <= rtol * abs(y)


def check_repetition(arr, nx: int, ny: int):
    """True if `arr` repeats with period nx along axis 1 and ny along axis 0 (the question
    hotspice/utils.py:42-66 answers): the (ny, nx) corner block, tiled over the array, matches
    every element within np.allclose tolerances."""
    arr = np.asarray(arr)
    H, W = arr.shape[:2]
    base = arr[:ny, :nx, ...]
    reps = (math.ceil(H / base.shape[0]), math.ceil(W / base.shape[1])) + (1,) * (arr.ndim - 2)
    return bool(np.allclose(np.tile(base, reps)[:H, :W, ...], arr))


def full_obj_name(obj):
    klass = type(obj)
    return f"{klass.__module__}.{klass.__qualname__}"


ELEMENTARY_CHARGE = 1.602176634e-19  # C (exact, SI 2019): joules per electronvolt


def eV_to_J(E, /):
    """utils.py:239-240"""
    return E * ELEMENTARY_CHARGE


def J_to_eV(E, /):
    """utils.py:241-242"""
    return E / ELEMENTARY_CHARGE


def asnumpy(array) -> np.ndarray:
    """utils.py:230-237: a NumPy array of `array` (host arrays pass through; torch tensors are
    copied to the host)."""
    if hasattr(array, "detach") and hasattr(array, "cpu"):
        return array.detach().cpu().numpy()
    return np.asarray(array)
