"""Host mirror of the reference CLbmSolver<T> operator surface over the C ABI.

``LbmSolverHIP`` keeps the reference's method names and argument meaning
(CLbmSolver.hpp:68-88, CLbmSolverGPU.cuh:56-97):

* boxes are (origin, size) in halo coordinates; ``None`` selects the
  reference's whole-domain overloads;
* host buffers are dense ``[component][z][y][x]`` numpy arrays (19 components
  for populations, 3 for velocities, 1 for densities and flags);
* device buffers may be any object exposing ``data_ptr()`` (e.g. a torch
  tensor on the GPU); then the copy is asynchronous on ``stream``.

Where the reference asserts or exits, this class raises ``LbmError``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LbmError, check, lib

_PRECISIONS = {"f32": _lib.F32, "f64": _lib.F64, "f16s": _lib.F16S, "bf16s": _lib.BF16S,
               "f16c": _lib.F16C}


def precision_code(precision) -> int:
    if isinstance(precision, str):
        return _PRECISIONS[precision.lower()]
    return int(precision)


def host_dtype(precision) -> np.dtype:
    return np.dtype(np.float64 if precision_code(precision) == _lib.F64 else np.float32)


def _i3(v):
    if v is None:
        return None
    return (ctypes.c_int * 3)(*[int(x) for x in v])


def _stream(stream):
    if stream is None:
        return None
    if hasattr(stream, "cuda_stream"):
        return ctypes.c_void_p(stream.cuda_stream)
    return ctypes.c_void_p(int(stream))


def _buffer(buf, dtype=None, writable=False):
    """(pointer, on_device, keepalive) of a numpy array or a device tensor."""
    if hasattr(buf, "data_ptr"):
        return ctypes.c_void_p(buf.data_ptr()), 1, buf
    arr = np.asarray(buf)
    if dtype is not None and arr.dtype != dtype:
        if writable:
            raise TypeError(f"host buffer must have dtype {dtype}, got {arr.dtype}")
        arr = arr.astype(dtype)
    if not arr.flags["C_CONTIGUOUS"]:
        if writable:
            raise TypeError("host buffer must be C-contiguous")
        arr = np.ascontiguousarray(arr)
    return arr.ctypes.data_as(ctypes.c_void_p), 0, arr


@dataclass
class SolverConfig:
    """Inputs of one subdomain solver (the CLbmSolver-derived quantities)."""
    size: tuple
    bc: tuple = (_lib.OBSTACLE,) * 6
    precision: str = "f32"
    inv_tau: float = 1.0
    lid_velocity: float = 0.0
    gravitation: tuple = (0.0, 0.0, 0.0)
    store_densities: bool = True
    store_velocities: bool = True
    device: int = 0
    extra: dict = field(default_factory=dict)


def make_params(size, bc=(_lib.OBSTACLE,) * 6, precision="f32", inv_tau=1.0, lid_velocity=0.0,
                gravitation=(0.0, 0.0, 0.0), store_densities=True, store_velocities=True,
                device=0):
    """The lbm_params struct of one subdomain solver."""
    p = _lib.lbm_params()
    p.size[:] = [int(v) for v in size]
    p.bc[:] = [int(v) for v in bc]
    p.precision = precision_code(precision)
    p.inv_tau = float(inv_tau)
    p.lid_velocity = float(lid_velocity)
    p.gravitation[:] = [float(v) for v in gravitation]
    p.store_densities = int(bool(store_densities))
    p.store_velocities = int(bool(store_velocities))
    p.device = int(device)
    return p


def device_bytes(size, precision="f32", store_outputs=False, pair_buffer=False) -> int:
    """Device bytes of one solver of this shape, residual snapshot included
    (and the second population buffer of the batch pair pass)."""
    p = make_params(size, precision=precision, store_densities=store_outputs,
                    store_velocities=store_outputs)
    b = ctypes.c_uint64()
    check(lib().lbm_device_bytes(ctypes.byref(p), int(bool(pair_buffer)), ctypes.byref(b)))
    return int(b.value)


class LbmSolverHIP:
    """One subdomain on one GPU (CLbmSolverGPU<T> replacement)."""

    def __init__(self, size, bc=(_lib.OBSTACLE,) * 6, precision="f32", inv_tau=1.0,
                 lid_velocity=0.0, gravitation=(0.0, 0.0, 0.0), store_densities=True,
                 store_velocities=True, device=0):
        L = lib()
        p = make_params(size, bc, precision, inv_tau, lid_velocity, gravitation,
                        store_densities, store_velocities, device)
        self._h = ctypes.c_void_p()
        check(L.lbm_create(ctypes.byref(p), ctypes.byref(self._h)))
        self.params = p
        self.size = tuple(int(v) for v in size)
        self.size_h = tuple(v + 2 for v in self.size)
        self.precision = p.precision
        self.dtype = host_dtype(p.precision)
        self.store_densities = bool(store_densities)
        self.store_velocities = bool(store_velocities)

    # -- lifecycle --------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lbm_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def device_arrays(self) -> _lib.lbm_device_arrays:
        out = _lib.lbm_device_arrays()
        check(lib().lbm_get_device_arrays(self._h, ctypes.byref(out)))
        return out

    def _cells(self, origin, size, interior=True):
        if origin is None:
            return int(np.prod(self.size if interior else self.size_h))
        return int(np.prod(size))

    # -- steps (CLbmSolver.hpp:68-71) ----------------------------------------
    def simulationStepAlpha(self, origin=None, size=None, stream=None):
        check(lib().lbm_simulation_step_alpha(self._h, _i3(origin), _i3(size), _stream(stream)))

    def simulationStepBeta(self, origin=None, size=None, stream=None):
        check(lib().lbm_simulation_step_beta(self._h, _i3(origin), _i3(size), _stream(stream)))

    # -- populations (CLbmSolver.hpp:73-76, CLbmSolverCPU.hpp:80) -----------
    def getDensityDistributions(self, origin=None, size=None, dst=None, stream=None):
        n = 19 * self._cells(origin, size)
        if dst is None:
            dst = np.empty(n, dtype=self.dtype)
        ptr, dev, keep = _buffer(dst, self.dtype, writable=True)
        check(lib().lbm_get_density_distributions(self._h, _i3(origin), _i3(size), ptr, dev,
                                                  _stream(stream)))
        return dst

    def setDensityDistributions(self, src, origin=None, size=None, direction=None,
                                stream=None):
        ptr, dev, keep = _buffer(src, self.dtype)
        if direction is None:
            check(lib().lbm_set_density_distributions(self._h, _i3(origin), _i3(size), ptr, dev,
                                                      _stream(stream)))
        else:
            check(lib().lbm_set_density_distributions_dir(self._h, _i3(origin), _i3(size),
                                                          int(direction), ptr, dev,
                                                          _stream(stream)))

    # -- flags (CLbmSolver.hpp:77-80) ---------------------------------------
    def getFlags(self, origin=None, size=None):
        dst = np.empty(self._cells(origin, size), dtype=np.int32)
        check(lib().lbm_get_flags(self._h, _i3(origin), _i3(size),
                                  dst.ctypes.data_as(ctypes.c_void_p)))
        return dst

    def setFlags(self, src, origin=None, size=None):
        arr = np.ascontiguousarray(src, dtype=np.int32)
        check(lib().lbm_set_flags(self._h, _i3(origin), _i3(size),
                                  arr.ctypes.data_as(ctypes.c_void_p)))

    # -- velocities / densities (CLbmSolver.hpp:81-88) ------------------------
    def getVelocities(self, origin=None, size=None, dst=None, stream=None):
        if dst is None:
            dst = np.empty(3 * self._cells(origin, size), dtype=self.dtype)
        ptr, dev, keep = _buffer(dst, self.dtype, writable=True)
        check(lib().lbm_get_velocities(self._h, _i3(origin), _i3(size), ptr, dev,
                                       _stream(stream)))
        return dst

    def setVelocities(self, src, origin=None, size=None, stream=None):
        ptr, dev, keep = _buffer(src, self.dtype)
        check(lib().lbm_set_velocities(self._h, _i3(origin), _i3(size), ptr, dev,
                                       _stream(stream)))

    def getDensities(self, origin=None, size=None, dst=None, stream=None):
        if dst is None:
            dst = np.empty(self._cells(origin, size), dtype=self.dtype)
        ptr, dev, keep = _buffer(dst, self.dtype, writable=True)
        check(lib().lbm_get_densities(self._h, _i3(origin), _i3(size), ptr, dev,
                                      _stream(stream)))
        return dst

    def setDensities(self, src, origin=None, size=None, stream=None):
        ptr, dev, keep = _buffer(src, self.dtype)
        check(lib().lbm_set_densities(self._h, _i3(origin), _i3(size), ptr, dev,
                                      _stream(stream)))

    def getMoments(self, layout: int, origin=None, size=None, dst=None, stream=None):
        """[rho, jx, jy, jz] per cell computed from the populations
        (layout 0: after beta/init, 1: after alpha)."""
        if dst is None:
            dst = np.empty(4 * self._cells(origin, size), dtype=self.dtype)
        ptr, dev, keep = _buffer(dst, self.dtype, writable=True)
        check(lib().lbm_get_moments(self._h, int(layout), _i3(origin), _i3(size), ptr, dev,
                                    _stream(stream)))
        return dst

    # -- halo faces (CComm shapes, CManager.cpp:135-202) ---------------------
    def face_box(self, direction, which):
        o = (ctypes.c_int * 3)()
        z = (ctypes.c_int * 3)()
        check(lib().lbm_face_box(self._h, int(direction), int(which), o, z))
        return tuple(o), tuple(z)

    def face_elements(self, direction) -> int:
        return int(lib().lbm_face_elements(self._h, int(direction)))

    def pack_face(self, direction, phase, dev_buf, stream=None):
        ptr, dev, keep = _buffer(dev_buf)
        if not dev:
            raise LbmError(_lib.LBM_OK - 1, "pack_face needs a device buffer")
        check(lib().lbm_pack_face(self._h, int(direction), int(phase), ptr, _stream(stream)))

    def unpack_face(self, direction, phase, dev_buf, stream=None):
        ptr, dev, keep = _buffer(dev_buf)
        if not dev:
            raise LbmError(_lib.LBM_OK - 1, "unpack_face needs a device buffer")
        check(lib().lbm_unpack_face(self._h, int(direction), int(phase), ptr, _stream(stream)))

    def exchange_face(self, direction, phase, to, stream=None):
        """lbm_exchange_face: pack_face(direction, phase) of this solver and
        unpack_face into `to` across the opposite face, in one launch with no
        buffer (two subdomains on one device)."""
        check(lib().lbm_exchange_face(self._h, int(direction), int(phase), to.handle,
                                      _stream(stream)))

    def _step_buffers(self, left, right, what):
        ptrs, keep = [], []
        for b in (left, right):
            if b is None:
                ptrs.append(None)
                continue
            ptr, dev, k = _buffer(b)
            if not dev:
                raise LbmError(_lib.LBM_OK - 1, f"{what} needs device buffers")
            ptrs.append(ptr)
            keep.append(k)
        return ptrs, keep

    def set_step_pack(self, left=None, right=None):
        """lbm_set_step_pack: the steps pack the LEFT / RIGHT x faces into
        these device buffers (None: that side off)."""
        (l, r), keep = self._step_buffers(left, right, "set_step_pack")
        check(lib().lbm_set_step_pack(self._h, l, r))
        self._step_pack_bufs = keep

    def set_step_unpack(self, left=None, right=None):
        """lbm_set_step_unpack: the steps take the x faces' received
        populations from these device buffers (None: that side off)."""
        (l, r), keep = self._step_buffers(left, right, "set_step_unpack")
        check(lib().lbm_set_step_unpack(self._h, l, r))
        self._step_unpack_bufs = keep

    def pack_face_pair(self, axis, phase, lo_buf, hi_buf, stream=None):
        """Both faces of one axis in one call: the low face (2 * axis) into
        lo_buf, the high face (2 * axis + 1) into hi_buf."""
        lo, lo_dev, keep_lo = _buffer(lo_buf)
        hi, hi_dev, keep_hi = _buffer(hi_buf)
        if not (lo_dev and hi_dev):
            raise LbmError(_lib.LBM_OK - 1, "pack_face_pair needs device buffers")
        check(lib().lbm_pack_face_pair(self._h, int(axis), int(phase), lo, hi, _stream(stream)))

    def unpack_face_pair(self, axis, phase, lo_buf, hi_buf, stream=None):
        """Both faces of one axis in one call: lo_buf into the low face
        (2 * axis), hi_buf into the high face (2 * axis + 1)."""
        lo, lo_dev, keep_lo = _buffer(lo_buf)
        hi, hi_dev, keep_hi = _buffer(hi_buf)
        if not (lo_dev and hi_dev):
            raise LbmError(_lib.LBM_OK - 1, "unpack_face_pair needs device buffers")
        check(lib().lbm_unpack_face_pair(self._h, int(axis), int(phase), lo, hi,
                                         _stream(stream)))

    # -- reductions -----------------------------------------------------------
    def reduce_box(self, origin=None, size=None, stream=None) -> np.ndarray:
        out = (ctypes.c_double * 5)()
        check(lib().lbm_reduce_box(self._h, _i3(origin), _i3(size), out, _stream(stream)))
        return np.array(out[:], dtype=np.float64)

    def residual(self, layout: int, stream=None) -> float:
        r = ctypes.c_double()
        check(lib().lbm_residual(self._h, int(layout), ctypes.byref(r), _stream(stream)))
        return r.value

    def sync(self, stream=None):
        check(lib().lbm_sync(self._h, _stream(stream)))

    def reset(self):
        """Populations back to the rest equilibrium (flags kept); clears the
        residual snapshot and the F16C overflow mark."""
        check(lib().lbm_reset(self._h))

    # -- checkpoint of the whole A-A state ---------------------------------
    def save_state(self) -> np.ndarray:
        """Populations of every cell (halo included, as stored) then flags,
        as one uint8 array; restore with load_state on a solver of the same
        size and precision."""
        n = ctypes.c_uint64()
        check(lib().lbm_state_bytes(self._h, ctypes.byref(n)))
        out = np.empty(n.value, dtype=np.uint8)
        check(lib().lbm_save_state(self._h, out.ctypes.data_as(ctypes.c_void_p), n.value))
        return out

    def load_state(self, state):
        arr = np.ascontiguousarray(state, dtype=np.uint8).ravel()
        check(lib().lbm_load_state(self._h, arr.ctypes.data_as(ctypes.c_void_p), arr.size))

    # -- CController::setDrivenCavitySzenario (CController.cpp:1313-1345) -----
    def setDrivenCavitySzenario(self):
        origin = (1, self.size_h[1] - 2, 1)
        size = (self.size[0], 1, self.size[2])
        self.setFlags(np.full(size[0] * size[2], _lib.VELOCITY_INJECTION, np.int32), origin,
                      size)


def face_transfers(transfers):
    """The lbm_face_transfer array of [(from_solver, direction, to_solver), ...]
    for exchange_faces; it holds the solvers' handles, so it is valid while
    they are open."""
    arr = (_lib.lbm_face_transfer * len(transfers))()
    for k, (a, d, b) in enumerate(transfers):
        arr[k].from_, arr[k].direction, arr[k].to = a.handle.value, int(d), b.handle.value
    return arr


def exchange_faces(transfers, phase, stream=None):
    """lbm_exchange_faces: the face transfers of one axis, a list of
    (from_solver, direction, to_solver) or a face_transfers array, in
    ceil(n / 8) launches, with the result of exchange_face on each in turn.
    All on one axis, device and precision, no two into the same face."""
    arr = transfers if isinstance(transfers, ctypes.Array) else face_transfers(transfers)
    check(lib().lbm_exchange_faces(arr, len(arr), int(phase), _stream(stream)))


def compute_parameters(length_x, size_x, length_max, velocity, acceleration=(0.0, 0.0, 0.0),
                       viscosity=-1.0, timestep=1.0, max_velocity=0.1, max_acceleration=0.1,
                       precision="f32") -> dict:
    """CLbmSolver constructor time-step adaptation (CLbmSolver.cpp:75-281)."""
    phys = _lib.lbm_physics()
    phys.length_x = float(length_x)
    phys.size_x = int(size_x)
    phys.length_max = float(length_max)
    phys.velocity[:] = [float(v) for v in velocity]
    phys.acceleration[:] = [float(v) for v in acceleration]
    phys.viscosity = float(viscosity)
    phys.timestep = float(timestep)
    phys.max_velocity = float(max_velocity)
    phys.max_acceleration = float(max_acceleration)
    out = _lib.lbm_derived()
    check(lib().lbm_compute_parameters(ctypes.byref(phys), precision_code(precision),
                                       ctypes.byref(out)))
    return dict(velocity_dimless=tuple(out.velocity_dimless),
                acceleration_dimless=tuple(out.acceleration_dimless),
                viscosity=out.viscosity, viscosity_dimless=out.viscosity_dimless, tau=out.tau,
                inv_tau=out.inv_tau, timestep=out.timestep)
