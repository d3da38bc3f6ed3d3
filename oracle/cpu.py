"""ctypes binding of liblbm_oracle.so (TEST INFRASTRUCTURE ONLY).

Arrays use the reference layout: dd[l * N_h + id], u[c * N_h + id], rho[id],
flags[id] (int32 Flag), id = (z * ny_h + y) * nx_h + x.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")

OBSTACLE, FLUID, VELOCITY_INJECTION, GHOST_LAYER = 1, 2, 4, 8
LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT = range(6)

_lib = None


def build(force: bool = False) -> None:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_i3 = ctypes.c_int * 3
_i6 = ctypes.c_int * 6
_P = ctypes.c_void_p


def _declare(L):
    L.oracle_set_flags.argtypes = [_P, _P, _P, ctypes.c_int]
    L.oracle_set_flags.restype = None
    L.oracle_direction_mask.argtypes = [ctypes.c_int]
    L.oracle_direction_mask.restype = ctypes.c_uint32
    for s, ct in (("f", ctypes.c_float), ("d", ctypes.c_double)):
        getattr(L, f"oracle_init_{s}").argtypes = [_P, _P, _P, _P, ctypes.c_int, ctypes.c_int]
        getattr(L, f"oracle_alpha_{s}").argtypes = [_P, _P, _P, _P, ct, ct, _P, _P, _P, _P,
                                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
        getattr(L, f"oracle_beta_{s}").argtypes = [_P, _P, _P, _P, ct, ct, _P, _P, _P,
                                                   ctypes.c_int, ctypes.c_int, ctypes.c_int]
        getattr(L, f"oracle_get_box_{s}").argtypes = [_P, ctypes.c_int, _P, _P, _P, _P]
        getattr(L, f"oracle_set_box_{s}").argtypes = [_P, ctypes.c_int, _P, _P, _P, _P,
                                                      ctypes.c_uint32]
        getattr(L, f"oracle_solver_params_{s}").argtypes = [ct, ctypes.c_int, ct, _P, _P, ct, ct,
                                                            ct, ct, _P, _P, _P, _P, _P, _P]
        getattr(L, f"oracle_solver_params_{s}").restype = ctypes.c_int
        getattr(L, f"oracle_moments_{s}").argtypes = [_P, _P, _P, _P, ctypes.c_int, _P]
        getattr(L, f"oracle_moments_{s}").restype = None
        for fn in ("init", "alpha", "beta", "get_box", "set_box"):
            getattr(L, f"oracle_{fn}_{s}").restype = None


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "f"
    if dtype == np.float64:
        return "d"
    raise TypeError(f"oracle supports float32/float64, not {dtype}")


def set_flags(size_h, bc, order: int = 0) -> np.ndarray:
    """CLbmInitCPU::setFlags; order 0 = x fastest, 1 = reference CPU z-fastest."""
    n = int(np.prod(size_h))
    flags = np.empty(n, dtype=np.int32)
    lib().oracle_set_flags(_ptr(flags), _i3(*size_h), _i6(*bc), int(order))
    return flags


def direction_mask(direction: int) -> int:
    return int(lib().oracle_direction_mask(int(direction)))


def solver_params(length_x, size_x, length_max, velocity, acceleration, viscosity, timestep,
                  max_velocity, max_acceleration, dtype=np.float32):
    """CLbmSolver constructor time-step adaptation. Returns (status, dict)."""
    s = _suffix(dtype)
    ct = ctypes.c_float if s == "f" else ctypes.c_double
    vel = np.asarray(velocity, dtype=dtype)
    acc = np.asarray(acceleration, dtype=dtype)
    ov = np.zeros(3, dtype=dtype)
    oa = np.zeros(3, dtype=dtype)
    outs = [ct() for _ in range(4)]
    st = getattr(lib(), f"oracle_solver_params_{s}")(
        ct(length_x), int(size_x), ct(length_max), _ptr(vel), _ptr(acc), ct(viscosity),
        ct(timestep), ct(max_velocity), ct(max_acceleration), _ptr(ov), _ptr(oa),
        *[ctypes.byref(o) for o in outs])
    return st, dict(velocity_dimless=ov, acceleration_dimless=oa,
                    viscosity_dimless=dtype(outs[0].value), tau=dtype(outs[1].value),
                    inv_tau=dtype(outs[2].value), timestep=dtype(outs[3].value))


class OracleSolver:
    """State of one CLbmSolverCPU-like subdomain held in numpy arrays."""

    def __init__(self, size, bc=(OBSTACLE,) * 6, dtype=np.float32, inv_tau=1.0,
                 lid_velocity=0.0, gravitation=(0.0, 0.0, 0.0), store_densities=True,
                 store_velocities=True, flag_order: int = 0, nthreads: int = 1):
        self.dtype = np.dtype(dtype)
        self.s = _suffix(self.dtype)
        self.size = tuple(int(v) for v in size)
        self.size_h = tuple(v + 2 for v in self.size)
        self.n = int(np.prod(self.size_h))
        self.inv_tau = self.dtype.type(inv_tau)
        self.lid = self.dtype.type(lid_velocity)
        self.g = np.asarray(gravitation, dtype=self.dtype)
        self.store_rho = bool(store_densities)
        self.store_u = bool(store_velocities)
        self.nthreads = int(nthreads)
        self.dd = np.empty(19 * self.n, dtype=self.dtype)
        self.rho = np.zeros(self.n, dtype=self.dtype) if self.store_rho else None
        self.u = np.zeros(3 * self.n, dtype=self.dtype) if self.store_u else None
        self.flags = set_flags(self.size_h, bc, flag_order)
        getattr(lib(), f"oracle_init_{self.s}")(_ptr(self.dd), _ptr(self.rho), _ptr(self.u),
                                               _i3(*self.size_h), int(self.store_rho),
                                               int(self.store_u))

    def _box(self, origin, size):
        if origin is None:
            return (0, 0, 0), self.size_h
        return tuple(origin), tuple(size)

    def alpha(self, origin=None, size=None):
        o, b = self._box(origin, size)
        getattr(lib(), f"oracle_alpha_{self.s}")(
            _ptr(self.dd), _ptr(self.flags), _ptr(self.rho), _ptr(self.u), self.inv_tau,
            self.lid, _ptr(self.g), _i3(*self.size_h), _i3(*o), _i3(*b), int(self.store_rho),
            int(self.store_u), self.nthreads)

    def beta(self, origin=None, size=None):
        o, b = self._box(origin, size)
        getattr(lib(), f"oracle_beta_{self.s}")(
            _ptr(self.dd), _ptr(self.flags), _ptr(self.rho), _ptr(self.u), self.inv_tau,
            self.lid, _i3(*self.size_h), _i3(*o), _i3(*b), int(self.store_rho),
            int(self.store_u), self.nthreads)

    def get_box(self, field: np.ndarray, ncomp: int, origin, size) -> np.ndarray:
        cells = int(np.prod(size))
        dst = np.empty(ncomp * cells, dtype=field.dtype)
        if field.dtype == np.int32:
            v = field.reshape(ncomp, self.size_h[2], self.size_h[1], self.size_h[0])
            (ox, oy, oz), (sx, sy, sz) = origin, size
            return np.ascontiguousarray(v[:, oz:oz + sz, oy:oy + sy, ox:ox + sx]).reshape(-1)
        getattr(lib(), f"oracle_get_box_{_suffix(field.dtype)}")(
            _ptr(field), ncomp, _i3(*self.size_h), _i3(*origin), _i3(*size), _ptr(dst))
        return dst

    def set_box(self, field: np.ndarray, ncomp: int, origin, size, src: np.ndarray,
                mask: int = 0xFFFFFFFF):
        src = np.ascontiguousarray(src, dtype=field.dtype)
        if field.dtype == np.int32:
            v = field.reshape(ncomp, self.size_h[2], self.size_h[1], self.size_h[0])
            (ox, oy, oz), (sx, sy, sz) = origin, size
            v[:, oz:oz + sz, oy:oy + sy, ox:ox + sx] = src.reshape(ncomp, sz, sy, sx)
            return
        getattr(lib(), f"oracle_set_box_{_suffix(field.dtype)}")(
            _ptr(field), ncomp, _i3(*self.size_h), _i3(*origin), _i3(*size), _ptr(src), mask)

    # CLbmSolver operator surface (box in halo coordinates, dense [c][z][y][x])
    def getDensityDistributions(self, origin, size):
        return self.get_box(self.dd, 19, origin, size)

    def setDensityDistributions(self, origin, size, src, direction=None):
        mask = 0x7FFFF if direction is None else direction_mask(direction)
        self.set_box(self.dd, 19, origin, size, src, mask)

    def getFlags(self, origin, size):
        return self.get_box(self.flags, 1, origin, size)

    def setFlags(self, origin, size, src):
        self.set_box(self.flags, 1, origin, size, np.asarray(src, dtype=np.int32))

    def getVelocities(self, origin=None, size=None):
        o, b = ((1, 1, 1), self.size) if origin is None else (origin, size)
        return self.get_box(self.u, 3, o, b)

    def getDensities(self, origin=None, size=None):
        o, b = ((1, 1, 1), self.size) if origin is None else (origin, size)
        return self.get_box(self.rho, 1, o, b)

    def moments(self, layout: int, origin=None, size=None) -> np.ndarray:
        """[rho, jx, jy, jz] of every cell of the box (default: interior) from
        the populations, in the reference accumulation order."""
        o, b = ((1, 1, 1), self.size) if origin is None else (origin, size)
        dst = np.empty(4 * int(np.prod(b)), dtype=self.dtype)
        getattr(lib(), f"oracle_moments_{self.s}")(_ptr(self.dd), _i3(*self.size_h), _i3(*o),
                                                  _i3(*b), int(layout), _ptr(dst))
        return dst

    def set_driven_cavity(self):
        """CController::setDrivenCavitySzenario (CController.cpp:1313-1345)."""
        origin = (1, self.size_h[1] - 2, 1)
        size = (self.size[0], 1, self.size[2])
        self.setFlags(origin, size, np.full(size[0] * size[2], VELOCITY_INJECTION, np.int32))
