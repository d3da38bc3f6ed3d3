"""Step driver of the lid-driven cavity: the CController / CManager caller
surface of the reference, on one GPU.

* ``cavity_parameters`` maps (edge, lid velocity, Re) to the CLbmSolver
  quantities with domainLength = domainSize (dx = 1) and timestep = 1, so the
  dimensionless lid velocity is u_lid and tau = 0.5 (6 nu + 1) with
  nu = u_lid * N / Re (CLbmSolver.cpp:144-281; Re as in :284, N the largest
  interior extent).
* ``CavityController`` owns one or more subdomain solvers laid out like
  CManager's rank grid (x fastest, CManager.cpp:34-58): OBSTACLE on global
  faces, GHOST_LAYER between subdomains (:91-104), the lid on the top layer of
  the top subdomains (:206-208, CController.cpp:1313-1345).
* ``computeNextStep`` alternates alpha / beta by step parity
  (CController.cpp:1278-1287). After each half-step the subdomains exchange
  faces on the device in the reference order x, then y, then z
  (CController.cpp:369-570 after alpha, 807-1008 after beta), with
  ``lbm_exchange_face`` standing in for the MPI send/recv pair; the transfers
  of one axis are independent, so ``lbm_exchange_faces`` runs each axis's in
  one launch.
* ``overlap=True`` keeps the reference's step order for that exchange
  (CController.cpp:209-309, 647-747): each subdomain first steps the two
  cells next to every communicating face (the halo and send planes), then its
  inner box, and the faces are exchanged on a second stream while the inner
  boxes run. The inner boxes touch neither the planes the exchange reads nor
  the slots it writes, so the result is bit-identical to the serial order.
* ``fused_x=True`` (a split along x only) exchanges the x faces inside the
  step kernels: each subdomain packs its x faces into its own buffers
  (``lbm_set_step_pack``) and its steps read the neighbours' packs as the
  received faces (``lbm_set_step_unpack``), so no x-face copy kernel runs.
  After alpha and after beta the packs carry disjoint planes, so one buffer
  per face serves both; the result is bit-identical to the serial order.
  With y or z splits as well the reference forwards received x values to the
  y / z neighbours' edge cells through their exchange, which would run before
  a fused unpack, so those decompositions keep ``lbm_exchange_face``.
* ``run`` times ``loops`` steps and reports the reference's benchmark figures
  (CController.cpp:1382-1490): GLUPS over interior cells and its
  "bandwidth" = GLUPS * (19 * 2 * sizeof(T) + sizeof(Flag) [+ 4 * sizeof(T)]).
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .solver import (LbmSolverHIP, compute_parameters, exchange_faces, face_transfers, host_dtype,
                     precision_code)


def cavity_parameters(size, u_lid, reynolds, precision="f32", max_velocity=0.1,
                      max_acceleration=0.1):
    """CLbmSolver-derived parameters of a cavity with dx = 1 and dt = 1."""
    n = max(size)
    viscosity = float(u_lid) * float(n) / float(reynolds)
    return compute_parameters(length_x=size[0], size_x=size[0], length_max=n,
                              velocity=(u_lid, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
                              viscosity=viscosity, timestep=1.0, max_velocity=max_velocity,
                              max_acceleration=max_acceleration, precision=precision)


class CavityController:
    """Global cavity of ``size`` interior cells split into ``split`` subdomains."""

    def __init__(self, size, u_lid=0.05, reynolds=100.0, precision="f32",
                 store_outputs=True, split=(1, 1, 1), device=0, params=None, stream=None,
                 overlap=False, xslab=2, fused_x=False):
        self.size = tuple(int(v) for v in size)
        self.split = tuple(int(v) for v in split)
        for a in range(3):
            if self.size[a] % self.split[a]:
                raise ValueError("domain size must be divisible by the split")
        self.sub = tuple(self.size[a] // self.split[a] for a in range(3))
        if fused_x and (overlap or self.split[0] < 2 or self.split[1] != 1 or self.split[2] != 1):
            raise ValueError("fused_x needs a split along x only, without overlap")
        self.precision = precision_code(precision)
        self.dtype = host_dtype(precision)
        self.params = params or cavity_parameters(self.size, u_lid, reynolds, precision)
        self.inv_tau = self.params["inv_tau"]
        self.lid = self.params["velocity_dimless"][0]
        self.store_outputs = bool(store_outputs)
        self.stream = stream
        self.simulationStepCounter = 0
        self.solvers = {}
        self.writers = {}
        self.global_writer = None
        self.config = None
        px, py, pz = self.split
        for k, j, i in itertools.product(range(pz), range(py), range(px)):
            bc = [_lib.GHOST_LAYER] * 6
            if i == 0:
                bc[0] = _lib.OBSTACLE
            if i == px - 1:
                bc[1] = _lib.OBSTACLE
            if j == 0:
                bc[2] = _lib.OBSTACLE
            if j == py - 1:
                bc[3] = _lib.OBSTACLE
            if k == 0:
                bc[4] = _lib.OBSTACLE
            if k == pz - 1:
                bc[5] = _lib.OBSTACLE
            s = LbmSolverHIP(self.sub, bc, precision, self.inv_tau, self.lid,
                             self.params["acceleration_dimless"], self.store_outputs,
                             self.store_outputs, device)
            if j == py - 1:
                s.setDrivenCavitySzenario()
            self.solvers[(i, j, k)] = s
        # the face transfers of each axis (both ways between every pair of
        # neighbours), handed to one batched exchange per axis
        self._transfers = []
        for axis in range(3):
            t = []
            for key, a in self.solvers.items():
                nb = list(key)
                nb[axis] += 1
                b = self.solvers.get(tuple(nb))
                if b is not None:
                    t.append((a, 2 * axis + 1, b))
                    t.append((b, 2 * axis, a))
            if t:
                self._transfers.append(face_transfers(t))
        self.overlap = bool(overlap) and len(self.solvers) > 1
        # x slabs may be thicker than the reference's two cells: a slab two cells
        # wide touches one cache line per row and plane for 8 bytes of data
        self.xslab = max(2, int(xslab))
        self._streams = []
        if self.overlap:
            from ._hip import Event, Stream
            for a in range(3):
                # two slabs of two cells must fit without overlapping
                if self.split[a] > 1 and self.sub[a] < 2:
                    raise ValueError("overlap needs at least 2 interior cells per subdomain "
                                     "along every split axis")
            if self.stream is None:
                self._streams.append(Stream(device))
                self.stream = self._streams[-1].handle
            self._exchange_stream = Stream(device)
            self._streams.append(self._exchange_stream)
            self._slabs_done = Event()
            self._exchange_done = Event()
            self._boxes = {key: self._slab_boxes(key) for key in self.solvers}
        self.device = device
        self.fused_x = bool(fused_x)
        self._xbufs = []
        self._x_recv = {}
        self._x_armed = False
        if self.fused_x:
            self._setup_fused_x()

    @classmethod
    def from_config(cls, cfg, device=0, stream=None, vis_format="netcdf") -> "CavityController":
        """The cavity an XML configuration describes (``config.LbmConfiguration``):
        its size, subdomain split, CLbmSolver parameters and outputs, with a
        writer per rank when visualization is switched on: netCDF, the
        reference's active writer (CController.cpp:120-135), "netcdf-global",
        its parallel build's one file for the whole cavity, or "vtk", the
        writer it keeps commented out there."""
        if vis_format not in ("netcdf", "netcdf-global", "vtk"):
            raise ValueError("vis_format must be 'netcdf', 'netcdf-global' or 'vtk', not "
                             f"{vis_format!r}")
        c = cls(cfg.domain_size, precision=cfg.precision, store_outputs=cfg.store_outputs,
                split=cfg.subdomains, device=device, params=cfg.solver_parameters(),
                stream=stream)
        c.config = cfg
        if cfg.do_visualization and vis_format == "netcdf-global":
            from .output import NetCdfGlobalWriter
            size, _, length = cfg.subdomain(0)
            c.global_writer = NetCdfGlobalWriter(cfg.visualization_rate, cfg.visualization_dir,
                                                 cfg.domain_size, size, length, cfg.precision)
        elif cfg.do_visualization:
            from .output import NetCdfWriter, VtkWriter
            px, py, _ = cfg.subdomains
            for (i, j, k) in c.solvers:
                rank = i + px * (j + py * k)
                size, origin, length = cfg.subdomain(rank)
                if vis_format == "vtk":
                    w = VtkWriter(rank, cfg.visualization_rate, cfg.visualization_dir, size,
                                  origin, length, cfg.precision)
                else:
                    w = NetCdfWriter(rank, cfg.visualization_rate, cfg.visualization_dir, size,
                                     length, cfg.precision)
                c.writers[(i, j, k)] = w
        return c

    def render(self, iteration: int) -> list:
        """CLbmVisualization::render of every rank (a file every `rate` iterations)."""
        if self.writers or self.global_writer:
            self.sync()
        if self.global_writer:
            p = self.global_writer.render(iteration, self)
            return [p] if p else []
        return [p for key, w in self.writers.items()
                if (p := w.render(iteration, self.solvers[key])) is not None]

    def _setup_fused_x(self):
        """One pack buffer per communicating x face; a subdomain's received
        LEFT face is its left neighbour's RIGHT pack and vice versa."""
        import ctypes
        L = lib()
        elem = np.dtype(self.dtype).itemsize
        pack = {}
        for key, s in self.solvers.items():
            bufs = [None, None]
            for d in (0, 1):
                nb = (key[0] + (1 if d else -1), key[1], key[2])
                if nb in self.solvers:
                    p = ctypes.c_void_p()
                    check(L.lbm_device_alloc(self.device, s.face_elements(d) * elem,
                                             ctypes.byref(p)))
                    self._xbufs.append(p)
                    bufs[d] = p
            pack[key] = bufs
            check(L.lbm_set_step_pack(s.handle, bufs[0], bufs[1]))
        for key in self.solvers:
            left = pack.get((key[0] - 1, key[1], key[2]))
            right = pack.get((key[0] + 1, key[1], key[2]))
            self._x_recv[key] = (left[1] if left else None, right[0] if right else None)

    def _arm_fused_x(self):
        # from the first beta on: before it nothing was exchanged, and the
        # first alpha reads the subdomains' own slots
        L = lib()
        for key, s in self.solvers.items():
            check(L.lbm_set_step_unpack(s.handle, *self._x_recv[key]))
        self._x_armed = True

    def close(self):
        for st in self._streams:
            st.close()
        self._streams = []
        if self._xbufs:
            L = lib()
            self.sync()
            for s in self.solvers.values():
                check(L.lbm_set_step_unpack(s.handle, None, None))
                check(L.lbm_set_step_pack(s.handle, None, None))
            for p in self._xbufs:
                L.lbm_device_free(self.device, p)
            self._xbufs = []
        if self.overlap:
            self._slabs_done.close()
            self._exchange_done.close()
            self.overlap = False
        # the transfer lists hold the solvers' raw handles
        self._transfers = []
        for s in self.solvers.values():
            s.close()
        self.solvers = {}

    # -- stepping ---------------------------------------------------------
    def _slab_boxes(self, key):
        """CController::stepAlpha's boxes (CController.cpp:209-249): for each
        face with a neighbour, in the order LEFT, RIGHT, BOTTOM, TOP, BACK,
        FRONT, the two cells next to it over the current inner box's extent
        of the other axes, the inner box shrinking by those two cells.
        Returns (slabs, inner or None when nothing is left inside)."""
        s = self.solvers[key]
        org, ext = [0, 0, 0], list(s.size_h)
        faces = [d for d in range(6)
                 if tuple(v + ((1 if d & 1 else -1) if a == d // 2 else 0)
                          for a, v in enumerate(key)) in self.solvers]
        nx_faces = sum(1 for d in faces if d < 2)
        # the x slab width, as long as the slabs do not meet
        wx = self.xslab if nx_faces * self.xslab <= s.size_h[0] else 2
        slabs = []
        for d in faces:
            axis, hi = d // 2, d & 1
            w = wx if axis == 0 else 2
            bo, bz = list(org), list(ext)
            bz[axis] = w
            if hi:
                bo[axis] = s.size_h[axis] - w
            else:
                org[axis] += w
            ext[axis] -= w
            slabs.append((tuple(bo), tuple(bz)))
        inner = (tuple(org), tuple(ext)) if min(ext) > 0 else None
        return slabs, inner

    def _step_overlapped(self, phase: int):
        step = "simulationStepBeta" if phase else "simulationStepAlpha"
        for key, s in self.solvers.items():
            for box in self._boxes[key][0]:
                getattr(s, step)(*box, stream=self.stream)
        self._slabs_done.record(self.stream)
        self._slabs_done.wait_on(self._exchange_stream.handle)
        self._exchange(phase, self._exchange_stream.handle)
        self._exchange_done.record(self._exchange_stream.handle)
        for key, s in self.solvers.items():
            inner = self._boxes[key][1]
            if inner is not None:
                getattr(s, step)(*inner, stream=self.stream)
        # the next half-step reads what the exchange wrote
        self._exchange_done.wait_on(self.stream)

    def _exchange(self, phase: int, stream=None):
        st = self.stream if stream is None else stream
        for arr in self._transfers:   # x, then y, then z
            exchange_faces(arr, phase, st)

    def stepAlpha(self):
        if self.overlap:
            return self._step_overlapped(0)
        for s in self.solvers.values():
            s.simulationStepAlpha(stream=self.stream)
        if self.fused_x:
            if not self._x_armed:
                self._arm_fused_x()
        elif len(self.solvers) > 1:
            self._exchange(0)

    def stepBeta(self):
        if self.overlap:
            return self._step_overlapped(1)
        if self.fused_x and not self._x_armed:
            raise RuntimeError("fused_x: the first step must be an alpha step")
        for s in self.solvers.values():
            s.simulationStepBeta(stream=self.stream)
        if not self.fused_x and len(self.solvers) > 1:
            self._exchange(1)

    def computeNextStep(self):
        if self.simulationStepCounter & 1:
            self.stepBeta()
        else:
            self.stepAlpha()
        self.simulationStepCounter += 1

    def sync(self):
        for s in self.solvers.values():
            s.sync(self.stream)
            break

    def run(self, loops: int) -> dict:
        """CController::run (CController.cpp:1401-1440): render(0), then
        each step followed by render(step counter) when writers are set."""
        if loops < 0:
            raise ValueError("loops must be >= 0 (the reference's endless run is not offered)")
        self.render(self.simulationStepCounter)
        self.sync()
        t0 = time.perf_counter()
        for _ in range(loops):
            self.computeNextStep()
            if self.writers or self.global_writer:
                self.render(self.simulationStepCounter)
        self.sync()
        elapsed = time.perf_counter() - t0
        cells = int(np.prod(self.size))
        t = self.dtype.itemsize
        used = 19 * 2 * t + 4 + (4 * t if self.store_outputs else 0)
        glups = loops * cells / elapsed * 1e-9
        return dict(loops=loops, time_s=elapsed, iterations_per_s=loops / elapsed,
                    glups=glups, mlups=glups * 1e3, bandwidth_gbs=glups * used)

    # -- extraction -----------------------------------------------------------
    def _gather(self, getter, ncomp, dtype=None):
        self.sync()   # the getters copy on the default stream
        out = np.empty((ncomp,) + self.size[::-1], dtype=dtype or self.dtype)
        sx, sy, sz = self.sub
        for (i, j, k), s in self.solvers.items():
            part = getattr(s, getter)().reshape(ncomp, sz, sy, sx)
            out[:, k * sz:(k + 1) * sz, j * sy:(j + 1) * sy, i * sx:(i + 1) * sx] = part
        return out

    def getVelocities(self) -> np.ndarray:
        """Interior velocities [3][z][y][x] of the whole cavity."""
        return self._gather("getVelocities", 3)

    def getDensities(self) -> np.ndarray:
        return self._gather("getDensities", 1)[0]

    def getFlags(self) -> np.ndarray:
        """Interior flags [z][y][x] of the whole cavity."""
        return self._gather("getFlags", 1, np.int32)[0]
