"""ctypes binding of the oracle's FP64 tangent (liboracle_fsi_tan.so).

TEST INFRASTRUCTURE ONLY: the checker of the device adjoint, imported by
tests/ and the adjoint leg of bench.py.  The library is fsi_oracle.c compiled
with a dual-number type for double (fsi_tangent.cpp): every array element is
a value followed by NDIR directional derivatives, and one step maps a state
with derivatives to the next state with derivatives (forward mode, the
switches of the primal frozen).

A dual field is held here as a C-contiguous array of shape (ny', nx', 1 + NDIR)
(the oracle's column-major (i, j) element order with the components
innermost); `pack` and `unpack` convert from and to Fortran (nx', ny') arrays.
"""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_void_p
from pathlib import Path

import numpy as np

from eulerianfsi_amd import _abi

from .oracle import OracleError, build_oracle

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle_fsi_tan.so"
FIELDS = ("U", "V", "phi", "resource", "sigma_xx", "sigma_yy", "sigma_xy")
_dp = POINTER(c_double)
_LIB = None


def load():
    global _LIB
    if _LIB is not None:
        return _LIB
    if not LIB_PATH.exists():
        build_oracle()
    lib = ctypes.CDLL(str(LIB_PATH))
    P = c_void_p
    sig = {
        "fsio_create": (c_int, [POINTER(_abi.Grid), POINTER(_abi.Bcs), POINTER(_abi.Models), POINTER(_abi.Numerics),
                                POINTER(P)]),
        "fsio_destroy": (None, [P]),
        "fsio_last_error": (c_char_p, [P]),
        "fsio_integrate_step": (c_int, [P] + [c_void_p] * 8 + [c_double, c_double, POINTER(_abi.StepReportC),
                                                               POINTER(_abi.StepStatsC)]),
        "fsio_eval_tendencies": (c_int, [P] + [c_void_p] * 6 + [c_double] + [c_void_p] * 8),
        "fsio_tangent_set_shear_modulus": (None, [c_double, _dp]),
        "fsio_tangent_ndir": (c_int, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def ndir() -> int:
    return load().fsio_tangent_ndir()


def pack(value, derivs=()):
    """Dual field from a Fortran (nx', ny') value and up to NDIR derivative arrays."""
    value = np.asarray(value, dtype=np.float64)
    nd = ndir()
    out = np.zeros((value.shape[1], value.shape[0], 1 + nd))
    out[:, :, 0] = value.T
    for k, d in enumerate(derivs):
        if d is not None:
            out[:, :, 1 + k] = np.asarray(d, dtype=np.float64).T
    return out


def unpack(a):
    """(value, [derivative arrays]) as Fortran (nx', ny') arrays."""
    return (np.asfortranarray(a[:, :, 0].T),
            [np.asfortranarray(a[:, :, 1 + k].T) for k in range(a.shape[2] - 1)])


class Tangent:
    """Oracle context whose step carries directional derivatives."""

    def __init__(self, metrics, bcs, models, numerics):
        self.lib = load()
        self.cfg = _abi.abi_config(metrics, bcs, models, numerics)
        g, b, m, n = self.cfg
        h = c_void_p()
        st = self.lib.fsio_create(ctypes.byref(g), ctypes.byref(b), ctypes.byref(m), ctypes.byref(n), ctypes.byref(h))
        self.h = h
        if st:
            raise OracleError(st, self.lib.fsio_last_error(h).decode())
        self.nx, self.ny = metrics.Nx, metrics.Ny
        self.shear_modulus = float(m.shear_modulus)
        self.correction = pack(np.zeros((self.nx, self.ny)))
        self.set_shear_modulus_derivatives(None)

    def close(self):
        if self.h:
            self.lib.fsio_destroy(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shear_modulus_derivatives(self, derivs):
        """Derivatives of the shear modulus along each direction (None: a constant)."""
        if derivs is None:
            self.lib.fsio_tangent_set_shear_modulus(self.shear_modulus, None)
            return
        d = np.zeros(ndir())
        d[:len(derivs)] = derivs
        self._gd = d
        self.lib.fsio_tangent_set_shear_modulus(self.shear_modulus, d.ctypes.data_as(_dp))

    def tendencies(self, s: dict, t=0.0) -> dict:
        """Stage tendencies (E_u, E_v, I_u, I_v, E_phi, E_sigma_*) of the dual
        state s (sigma after the safety reset, as the oracle's entry point
        applies it), as dual fields."""
        nx, ny = self.nx, self.ny
        nd = ndir() + 1
        shapes = dict(E_u=(ny, nx + 1), E_v=(ny + 1, nx), I_u=(ny, nx + 1), I_v=(ny + 1, nx), E_phi=(ny, nx),
                      E_sigma_xx=(ny, nx), E_sigma_yy=(ny, nx), E_sigma_xy=(ny, nx))
        out = {k: np.zeros(v + (nd,)) for k, v in shapes.items()}
        args = [s[k].ctypes.data for k in ("U", "V", "phi", "sigma_xx", "sigma_yy", "sigma_xy")]
        st = self.lib.fsio_eval_tendencies(self.h, *args, float(t), *(out[k].ctypes.data for k in shapes))
        if st:
            raise OracleError(st, self.lib.fsio_last_error(self.h).decode())
        return out

    def integrate_step(self, s: dict, t, dt):
        """s: dict of dual fields (pack) for FIELDS, updated in place."""
        for k in FIELDS:
            assert s[k].flags["C_CONTIGUOUS"] and s[k].dtype == np.float64
        rep, stats = _abi.StepReportC(), _abi.StepStatsC()
        args = [s[k].ctypes.data for k in FIELDS] + [self.correction.ctypes.data]
        st = self.lib.fsio_integrate_step(self.h, *args, float(t), float(dt), ctypes.byref(rep), ctypes.byref(stats))
        if st:
            raise OracleError(st, self.lib.fsio_last_error(self.h).decode())
        return rep, stats
