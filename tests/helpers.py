"""Shared helpers of the test suite (not a test module)."""
import numpy as np


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    den = np.linalg.norm(b)
    num = np.linalg.norm(a - b)
    if den == 0.0:
        return num
    return num / den


def compact_u(Ufull):
    """(Nx+1, Ny) x-face field -> compact (Nx, Ny) u part."""
    return np.asarray(Ufull)[:-1, :]


def compact_v(Vfull):
    return np.asarray(Vfull)[:, :-1]
