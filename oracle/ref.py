"""ctypes binding of _ref/libref_lbm.so (TEST INFRASTRUCTURE ONLY).

That library is the reference's own CPU kernels (CLbmInitCPU, CLbmAlphaCPU,
CLbmBetaCPU, CLbmSolver constructor) compiled from /root/reference/src by
oracle/Makefile, wrapped by oracle/ref_harness.cpp. It is used to write the
golden vectors and, where present, as the timed CPU baseline. ``available()``
is False where it was not built (no reference source tree).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libref_lbm.so")
REF_ROOT = os.environ.get("REF", "/root/reference")   # the reference tree (fixture source)

_lib = None
_i3 = ctypes.c_int * 3
_i6 = ctypes.c_int * 6
_P = ctypes.c_void_p


def available() -> bool:
    return os.path.exists(LIB_PATH)


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB_PATH)
        for s, ct in (("f", ctypes.c_float), ("d", ctypes.c_double)):
            getattr(_lib, f"ref_init_{s}").argtypes = [_P, _P, _P, _P, _P, _P, ctypes.c_int,
                                                      ctypes.c_int]
            getattr(_lib, f"ref_alpha_{s}").argtypes = [_P, _P, _P, _P, ct, ct, _P, _P, _P, _P,
                                                       _P, ctypes.c_int, ctypes.c_int]
            getattr(_lib, f"ref_beta_{s}").argtypes = [_P, _P, _P, _P, ct, ct, _P, _P, _P, _P,
                                                      ctypes.c_int, ctypes.c_int]
            getattr(_lib, f"ref_solver_params_{s}").argtypes = [ctypes.c_char_p, _P]
            getattr(_lib, f"ref_vtk_{s}").argtypes = [ctypes.c_char_p, ctypes.c_char_p,
                                                     ctypes.c_int, ctypes.c_int, _P, _P, _P,
                                                     _P, _P, _P]
            getattr(_lib, f"ref_config_{s}").argtypes = [ctypes.c_char_p, _P, ctypes.c_char_p,
                                                        ctypes.c_int]
    return _lib


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _s(dtype):
    return "f" if np.dtype(dtype) == np.float32 else "d"


class RefSolver:
    """Arrays + calls of the reference CPU kernels, CLbmSolverCPU-style."""

    def __init__(self, size, bc=(1,) * 6, dtype=np.float32, inv_tau=1.0, lid_velocity=0.0,
                 gravitation=(0.0, 0.0, 0.0), store_densities=True, store_velocities=True,
                 block=(32, 32, 32)):
        self.dtype = np.dtype(dtype)
        self.s = _s(dtype)
        self.size = tuple(int(v) for v in size)
        self.size_h = tuple(v + 2 for v in self.size)
        self.n = int(np.prod(self.size_h))
        self.inv_tau = self.dtype.type(inv_tau)
        self.lid = self.dtype.type(lid_velocity)
        self.g = np.asarray(gravitation, dtype=self.dtype)
        self.store_rho = bool(store_densities)
        self.store_u = bool(store_velocities)
        self.block = tuple(block)
        self.dd = np.empty(19 * self.n, dtype=self.dtype)
        self.flags = np.empty(self.n, dtype=np.int32)
        self.rho = np.zeros(self.n, dtype=self.dtype) if self.store_rho else None
        self.u = np.zeros(3 * self.n, dtype=self.dtype) if self.store_u else None
        getattr(lib(), f"ref_init_{self.s}")(_ptr(self.dd), _ptr(self.flags), _ptr(self.rho),
                                            _ptr(self.u), _i3(*self.size_h), _i6(*bc),
                                            int(self.store_rho), int(self.store_u))

    def _box(self, origin, size):
        if origin is None:
            return (0, 0, 0), self.size_h
        return tuple(origin), tuple(size)

    def alpha(self, origin=None, size=None):
        o, b = self._box(origin, size)
        getattr(lib(), f"ref_alpha_{self.s}")(
            _ptr(self.dd), _ptr(self.flags), _ptr(self.rho), _ptr(self.u), self.inv_tau,
            self.lid, _ptr(self.g), _i3(*self.size_h), _i3(*o), _i3(*b), _i3(*self.block),
            int(self.store_rho), int(self.store_u))

    def beta(self, origin=None, size=None):
        o, b = self._box(origin, size)
        getattr(lib(), f"ref_beta_{self.s}")(
            _ptr(self.dd), _ptr(self.flags), _ptr(self.rho), _ptr(self.u), self.inv_tau,
            self.lid, _i3(*self.size_h), _i3(*o), _i3(*b), _i3(*self.block),
            int(self.store_rho), int(self.store_u))


def solver_params(xml_path: str, dtype=np.float32) -> np.ndarray:
    """CLbmSolver constructor on a configuration file: [u*(3), g*(3), nu*, tau,
    tauInv, timestep, viscosity, 0]. The reference exits the process on an
    infeasible configuration, so call this in a child process for such cases."""
    out = np.zeros(12, dtype=dtype)
    getattr(lib(), f"ref_solver_params_{_s(dtype)}")(xml_path.encode(), _ptr(out))
    return out


def vtk(xml_path, out_dir, rank, iteration, size, origin, length, flags, rho, u, dtype):
    """CLbmVisualizationVTK<T>::render of the given interior fields (flags
    [z][y][x], rho [z][y][x], u [3][z][y][x]) into out_dir."""
    dtype = np.dtype(dtype)
    f = np.ascontiguousarray(flags, dtype=np.int32)
    r = np.ascontiguousarray(rho, dtype=dtype)
    v = np.ascontiguousarray(u, dtype=dtype)
    ln = np.array(length, dtype=dtype)
    getattr(lib(), f"ref_vtk_{_s(dtype)}")(xml_path.encode(), out_dir.encode(), int(rank),
                                           int(iteration), _ptr(np.array(size, np.int32)),
                                           _ptr(np.array(origin, np.int32)), _ptr(ln), _ptr(f),
                                           _ptr(r), _ptr(v))


def config(xml_path, dtype):
    """Every field CConfiguration<T> reads (44 doubles, see ref_harness.cpp)
    and the visualization output directory."""
    out = np.zeros(44, dtype=np.float64)
    buf = ctypes.create_string_buffer(4096)
    getattr(lib(), f"ref_config_{_s(dtype)}")(xml_path.encode(), _ptr(out), buf, 4096)
    return out, buf.value.decode()
