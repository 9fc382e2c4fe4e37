"""Small helpers mirroring lagrbfpy's utilities used on the interpolation path.

PointContainer (utils/points_container.py), FunctionData / FunctionPointer
(interface/function_pointer.py:32-120), the Halton sequence (utils/halton_seq.py:39-78)
and the structured / Halton point generators (swarm_utils.py:159-273).
"""
import ctypes
from dataclasses import dataclass

import numpy as np


@dataclass
class PointContainer:
    coords: np.ndarray
    values: np.ndarray = None
    pidx: np.ndarray = None


class FunctionData(ctypes.Structure):
    """User data passed to velocity callbacks (time and step fields, as the reference)."""
    _fields_ = [("time", ctypes.c_double), ("step", ctypes.c_int32)]


class FunctionPointer:
    """Holds a Python batch callback f(npoints, dim, coords, velocity, data_ptr) -> int as a C pointer."""

    def __init__(self, func, data=None):
        from ._lib import VELOCITY_FN_BATCH
        if data is not None and not isinstance(data, ctypes.Structure):
            raise TypeError("data must be a subclass of ctypes.Structure or FunctionData.")
        if not callable(func):
            raise TypeError("func must be callable with signature (npoints, dim, coords, velocity, data_ptr)")
        self.func = func
        self.data = data
        self.cfunc = VELOCITY_FN_BATCH(func)


def numpy_velocity_callback(pyfunc):
    """Wrap pyfunc(X[n,dim]) -> V[n,dim] as a batch callback."""
    def cb(n, dim, X, V, user):
        x = np.ctypeslib.as_array(X, shape=(n, dim))
        v = np.ctypeslib.as_array(V, shape=(n, dim))
        v[:, :] = pyfunc(x)
        return 0
    return cb


def halton_sequence(index, base):
    """Radical inverse of index in the given base (halton_seq.py:39-78)."""
    if base <= 1:
        raise ValueError("base must be greater than 1.")
    if isinstance(index, (int, np.integer)):
        f, r, i = 1.0, 0.0, int(index)
        while i > 0:
            f = f / float(base)
            r = r + f * float(i % base)
            i = i // base
        return r
    idx = np.asarray(index, dtype=np.int64).copy()
    f = np.ones_like(idx, dtype=np.float64)
    r = np.zeros_like(idx, dtype=np.float64)
    while np.any(idx > 0):
        f /= float(base)
        r += f * (idx % base)
        idx //= base
    return r


def structured_points(O, L, n):
    """Uniform grid, x fastest (swarm_utils.py:159-175: meshgrid 'ij' raveled in Fortran order)."""
    O = np.asarray(O, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    dim = O.shape[0]
    x = [np.linspace(O[d], L[d], int(n[d])) for d in range(dim)]
    X = np.meshgrid(*x, indexing="ij")
    return np.stack([X[d].ravel(order="F") for d in range(dim)], axis=1)


def halton_points(O, L, n, bases=(2, 3, 5)):
    """Halton points with indices 1..n (swarm_utils.py:228-273)."""
    O = np.asarray(O, dtype=np.float64)
    L = np.asarray(L, dtype=np.float64)
    idx = np.arange(int(n), dtype=np.int64) + 1
    h = np.stack([halton_sequence(idx, bases[d]) for d in range(O.shape[0])], axis=1)
    return O + h * (L - O)
