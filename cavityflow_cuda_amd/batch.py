"""Many independent lid-driven cavities on one GPU.

``CavityBatch`` owns one ``LbmSolverHIP`` per cavity (own arrays, own lid, own
tau and lid velocity) and steps them all with one launch per half-step through
``lbm_batch_*``; runs of steps are replayed from a captured HIP graph. Because
every member runs the same per-cell code as a lone solver, its result is
bit-identical alone and in the batch.

``draw_cavities`` produces seeded cavity parameters. Parameters go through the
reference's time-step adaptation (CLbmSolver.cpp:121-281, via
``cavity_parameters``); a draw it rejects (the reference would exit) is redrawn
and counted.
"""
from __future__ import annotations

import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LbmError, check, lib
from .controller import cavity_parameters
from .solver import LbmSolverHIP, _stream


@dataclass
class CavitySpec:
    size: tuple
    u_lid: float
    reynolds: float
    params: dict = field(default_factory=dict)   # CLbmSolver-derived (tau, lid in lattice units)

    @property
    def cells(self) -> int:
        return int(np.prod(self.size))


def draw_cavities(count, seed, edge=(32, 192), max_aspect=2.0, reynolds=(100.0, 2000.0),
                  u_lid=(0.02, 0.1), precision="f32"):
    """Seeded draws: every edge log-uniform in `edge` with longest/shortest <=
    max_aspect, Re and lid velocity uniform. Returns (specs, rejected), where
    `rejected` counts draws the reference adaptation cannot stabilise."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(edge[0]), math.log(edge[1])
    specs, rejected = [], 0
    while len(specs) < count:
        size = tuple(int(round(math.exp(rng.uniform(lo, hi)))) for _ in range(3))
        if max(size) > max_aspect * min(size):
            continue
        re = float(rng.uniform(*reynolds))
        u = float(rng.uniform(*u_lid))
        try:
            params = cavity_parameters(size, u, re, precision)
        except LbmError:
            rejected += 1
            continue
        specs.append(CavitySpec(size, u, re, params))
    return specs, rejected


class CavityBatch:
    """Cavities stepped together; all members stay at the same step parity."""

    def __init__(self, specs, precision="f32", device=0, store_outputs=False):
        self.specs = list(specs)
        self.precision = precision
        self.solvers = []
        for sp in self.specs:
            p = sp.params or cavity_parameters(sp.size, sp.u_lid, sp.reynolds, precision)
            sp.params = p
            s = LbmSolverHIP(sp.size, (_lib.OBSTACLE,) * 6, precision, p["inv_tau"],
                             p["velocity_dimless"][0], p["acceleration_dimless"],
                             store_outputs, store_outputs, device)
            s.setDrivenCavitySzenario()
            self.solvers.append(s)
        arr = (ctypes.c_void_p * len(self.solvers))(*[s.handle.value for s in self.solvers])
        self._h = ctypes.c_void_p()
        check(lib().lbm_batch_create(arr, len(self.solvers), ctypes.byref(self._h)))
        self.steps = 0

    @property
    def cells(self) -> int:
        return sum(sp.cells for sp in self.specs)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().lbm_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        for s in self.solvers:
            s.close()
        self.solvers = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- tuning -------------------------------------------------------------
    def tuning(self) -> dict:
        t = _lib.lbm_batch_tuning()
        check(lib().lbm_batch_get_tuning(self._h, ctypes.byref(t)))
        return dict(block=t.block, cells_per_thread=t.cells_per_thread,
                    waves_per_cu=t.waves_per_cu, options=(t.options[0], t.options[1]),
                    pairs=t.pairs, group_mb=t.group_mb)

    def set_tuning(self, block=256, cells_per_thread=1, waves_per_cu=0, options=None, pairs=0,
                   group_mb=0):
        """Launch shape; options = (alpha, beta) launch-option bits, None keeps
        the current ones (initially the precision's defaults); pairs=1 runs
        each alpha + beta pair as one out-of-place pass (then run() needs
        multiples of 4 steps); group_mb > 0 runs the members in groups of at
        most that many MiB of populations, one group after the other."""
        if options is None:
            options = self.tuning()["options"]
        t = _lib.lbm_batch_tuning(block, cells_per_thread, waves_per_cu,
                                  (ctypes.c_int * 2)(*options), int(pairs), int(group_mb))
        check(lib().lbm_batch_set_tuning(self._h, ctypes.byref(t)))

    def autotune(self, trial_steps=3, stream=None) -> dict:
        """Pick the fastest launch shape on this batch, then reset every member
        to its initial state (the trials advance the simulation)."""
        t = _lib.lbm_batch_tuning()
        check(lib().lbm_batch_autotune(self._h, int(trial_steps), _stream(stream), ctypes.byref(t)))
        self.reset()
        return self.tuning()

    def reset(self):
        for s in self.solvers:
            check(lib().lbm_reset(s.handle))
        self.steps = 0

    # -- stepping -------------------------------------------------------------
    def run(self, steps: int, stream=None, chunk: int = 200, graph: bool = True):
        """`steps` time steps (alternating alpha and beta, as
        CController::computeNextStep counts them; even, so every member ends
        in the natural layout), replayed from a captured graph of `chunk`
        steps on a non-default stream (graph=False: the same launches issued
        directly)."""
        if steps % 2 or chunk % 2 or chunk <= 0:
            raise ValueError("steps and chunk must be even (whole alpha + beta pairs)")
        st = _stream(stream)
        fn = lib().lbm_batch_run if graph else lib().lbm_batch_launch
        full, rest = divmod(int(steps), chunk)
        for _ in range(full):
            check(fn(self._h, 0, chunk, st))
        if rest:
            check(fn(self._h, 0, rest, st))
        self.steps += int(steps)

    def sync(self, stream=None):
        self.solvers[0].sync(stream)

    def residuals(self, stream=None) -> np.ndarray:
        out = (ctypes.c_double * len(self.solvers))()
        check(lib().lbm_batch_residual(self._h, 0, out, _stream(stream)))
        return np.array(out[:])

    def moments(self, i: int) -> np.ndarray:
        """[rho, jx, jy, jz] over the interior of member i (after whole steps)."""
        return self.solvers[i].getMoments(0)


def solve_alone(spec, steps, precision="f32", device=0, stream=None):
    """One cavity on its own solver for `steps` (even) time steps: one launch
    per step, no batching, no graph (the plain sequential baseline).
    Returns (moments, seconds)."""
    if steps % 2:
        raise ValueError("steps must be even (whole alpha + beta pairs)")
    p = spec.params or cavity_parameters(spec.size, spec.u_lid, spec.reynolds, precision)
    s = LbmSolverHIP(spec.size, (_lib.OBSTACLE,) * 6, precision, p["inv_tau"],
                     p["velocity_dimless"][0], p["acceleration_dimless"], False, False, device)
    s.setDrivenCavitySzenario()
    L = lib()
    st = _stream(stream)
    s.sync(stream)
    t0 = time.perf_counter()
    for _ in range(int(steps) // 2):
        check(L.lbm_simulation_step_alpha(s.handle, None, None, st))
        check(L.lbm_simulation_step_beta(s.handle, None, None, st))
    s.sync(stream)
    dt = time.perf_counter() - t0
    m = s.getMoments(0)
    s.close()
    return m, dt
