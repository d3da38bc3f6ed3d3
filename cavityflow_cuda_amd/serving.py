"""Serving lid-driven cavity solves on one GPU with continuous batching.

A request is one cavity (interior size, lid velocity, Reynolds number) to be
run until the steady-state residual ||u^n - u^(n-K)||_2 / ||u^n||_2 (FP64,
FLUID cells, K = ``check_every`` time steps, even) falls to ``tol``, or to a
step cap. A time step is one alpha or one beta half of the A-A pair, as
CController::computeNextStep counts them.
Its CLbmSolver parameters come from the reference's time-step adaptation
(CLbmSolver.cpp:121-281); a request the adaptation cannot stabilise (the
reference would exit) is rejected.

``CavityServer`` keeps every admitted cavity resident on the device as its own
solver and steps them in batches: each scheduling round replays one captured
graph of K time steps over the current batch (one launch per half-step
covering every member, ``lbm_batch_run``), then reads all its residuals with
one launch (``lbm_batch_residual``). Finished members leave and queued
requests join between rounds -- every member is at an even step there, so
the batch always runs alpha first. Which resident solves a round advances is
the policy: "srpt" takes them in order of least remaining work (cells x steps
left to the cap) until the batch holds ``min_batch_cells`` cells, enough to
fill the GPU, so short solves are not held behind long ones; "srpt_oldest"
adds the longest-resident solve to that batch every round, so no solve
waits behind an endless stream of shorter ones; "hrrn" (highest response
ratio next) orders them by (time since arrival + remaining work at the
measured rate) / remaining work instead, so waiting raises a long solve's
priority steadily (the default: at 80 % load it measured the lowest p50 and
p99 of the three); "all" advances every resident solve every round
(processor sharing). Admission is FIFO
under an HBM budget: a request is admitted when its device footprint
(``lbm_device_bytes``) fits in what the resident solves leave free; one that
can never fit is rejected.

``solve_sequential`` is the plain baseline: requests one after another in
arrival order, each on its own solver with one launch per half-step and the
same residual test.

Each member runs the same per-cell code as a lone solver, and the batch
residual of a member carries the same bits as its own lbm_residual, so a
request stops at the same step and returns the same bits whether it ran
alone or in a mix.
"""
from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import LbmError, check, lib
from .controller import cavity_parameters
from .solver import LbmSolverHIP, _stream, device_bytes


_ELEM_BYTES = {"f32": 4, "f64": 8, "f16s": 2, "bf16s": 2, "f16c": 2}


@dataclass
class SolveRequest:
    id: int
    size: tuple
    u_lid: float
    reynolds: float
    arrival: float = 0.0            # seconds after the serving clock starts
    tol: float = 1e-6
    max_steps: int = 50000
    keep_fields: bool = False       # return the final [rho, jx, jy, jz] interior fields
    initial: np.ndarray | None = None   # optional start: LbmSolverHIP.save_state() (a
    #                                     checkpoint, at an even step) or interior
    #                                     populations [19][z][y][x]
    params: dict | None = None      # CLbmSolver-derived (filled on submit)

    @property
    def cells(self) -> int:
        return int(np.prod(self.size))


@dataclass
class SolveResult:
    id: int
    status: str                     # converged | step_cap | failed | rejected | cancelled
    steps: int = 0
    residual: float = math.inf
    arrival: float = 0.0
    admitted: float = math.nan
    finished: float = math.nan
    totals: np.ndarray | None = None    # FP64 [sum f, sum c f (3), cells] over the interior
    fields: np.ndarray | None = None
    message: str = ""

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def draw_requests(count, seed, edge=(48, 384), max_aspect=2.0, reynolds=(100.0, 3200.0),
                  u_lid=0.1, rate=None, tol=1e-6, max_steps=50000, precision="f32"):
    """Seeded requests: edges log-uniform in `edge` (longest <= max_aspect x
    shortest), Re uniform, fixed lid velocity, Poisson arrivals at `rate` per
    second (all at 0 when rate is None; the arrival times come from their own
    stream, so the cavities do not depend on the rate). Draws the reference
    adaptation rejects are redrawn; returns (requests, redrawn)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(edge[0]), math.log(edge[1])
    out, redrawn = [], 0
    while len(out) < count:
        size = tuple(int(round(math.exp(rng.uniform(lo, hi)))) for _ in range(3))
        if max(size) > max_aspect * min(size):
            continue
        re = float(rng.uniform(*reynolds))
        try:
            params = cavity_parameters(size, u_lid, re, precision)
        except LbmError:
            redrawn += 1
            continue
        out.append(SolveRequest(len(out), size, float(u_lid), re, 0.0, tol, max_steps,
                                params=params))
    if rate:
        set_poisson_arrivals(out, rate, seed)
    return out, redrawn


def set_poisson_arrivals(requests, rate, seed):
    """Arrival times of a Poisson stream at `rate` per second, in list order."""
    gaps = np.random.default_rng([seed, 1]).exponential(1.0 / rate, len(requests))
    for r, t in zip(requests, np.cumsum(gaps)):
        r.arrival = float(t)


def _new_solver(req, precision, device):
    p = req.params
    s = LbmSolverHIP(req.size, (_lib.OBSTACLE,) * 6, precision, p["inv_tau"],
                     p["velocity_dimless"][0], p["acceleration_dimless"], False, False, device)
    s.setDrivenCavitySzenario()
    if req.initial is not None:
        if req.initial.dtype == np.uint8:
            s.load_state(req.initial)
        else:
            s.setDensityDistributions(np.ascontiguousarray(req.initial, dtype=s.dtype).ravel())
    return s


def _finish(req, s, status, steps, residual, stream):
    """Result of a solve whose populations are in the natural (post-beta) layout."""
    r = SolveResult(req.id, status, steps, residual, req.arrival)
    inner, size = (1, 1, 1), tuple(req.size)
    r.totals = s.reduce_box(inner, size, stream)
    if req.keep_fields:
        r.fields = s.getMoments(0, stream=stream).reshape(4, *size[::-1])
    return r


class CavityServer:
    """Continuous-batching solve server on one GPU (see module docstring)."""

    def __init__(self, precision="f32", hbm_budget=200e9, check_every=100, device=0,
                 stream=None, tuning=None, max_members=4096, policy="hrrn",
                 min_batch_cells=16_000_000, graph=True):
        if check_every % 2 or check_every <= 0:
            raise ValueError("check_every must be even (whole alpha + beta pairs)")
        self.precision = precision
        self.budget = int(hbm_budget)
        self.K = int(check_every)
        self.device = int(device)
        self.stream = stream
        self.tuning = tuning
        self.max_members = int(max_members)
        if policy not in ("srpt", "srpt_oldest", "hrrn", "all"):
            raise ValueError("policy is 'srpt', 'srpt_oldest', 'hrrn' or 'all'")
        self.policy = policy
        self.min_batch_cells = int(min_batch_cells)
        self.graph = bool(graph)        # False: the round's launches issued directly
        self._batch_ids: tuple = ()
        self.queue: deque = deque()
        self.active: list = []          # [req, solver, bytes, steps, admitted]
        self.used = 0
        self.peak_used = 0
        self.results: dict = {}
        self.cell_updates = 0
        self.rounds = 0
        self._batch = ctypes.c_void_p()
        self._dirty = True
        self._cancel: set = set()
        self.t0 = time.perf_counter()

    def clock(self) -> float:
        return time.perf_counter() - self.t0

    # -- requests -------------------------------------------------------------
    def submit(self, req: SolveRequest):
        """Validate and queue; rejects at once what can never run."""
        if req.max_steps % self.K or req.max_steps <= 0:
            self._reject(req, f"max_steps must be a positive multiple of {self.K}")
            return
        if req.params is None:
            try:
                req.params = cavity_parameters(req.size, req.u_lid, req.reynolds, self.precision)
            except LbmError as e:
                self._reject(req, str(e))
                return
        try:
            req._bytes = device_bytes(req.size, self.precision,
                                      pair_buffer=bool(self.tuning and self.tuning.get("pairs")))
        except LbmError as e:
            self._reject(req, str(e))
            return
        if req._bytes > self.budget:
            self._reject(req, f"needs {req._bytes} B, over the {self.budget} B budget")
            return
        # batch launches address a plane with 32-bit byte offsets
        cells = int(np.prod([int(v) + 2 for v in req.size]))
        plane = (cells + 255) // 256 * 256 * _ELEM_BYTES[self.precision]
        if plane >= 1 << 32:
            self._reject(req, f"a {plane} B population plane is too large for a batch member")
            return
        self.queue.append(req)

    def cancel(self, req_id: int):
        """Drop a queued or running request at the next round boundary."""
        self._cancel.add(req_id)

    def _reject(self, req, msg):
        self.results[req.id] = SolveResult(req.id, "rejected", arrival=req.arrival,
                                           finished=req.arrival, message=msg)

    # -- scheduling ------------------------------------------------------------
    def _admit(self, now):
        admitted = 0
        while self.queue and self.queue[0].arrival <= now and len(self.active) < self.max_members:
            req = self.queue[0]
            if self.used + req._bytes > self.budget:
                break                                   # FIFO: wait for memory
            try:
                s = _new_solver(req, self.precision, self.device)
            except LbmError as e:
                if e.status == _lib.LBM_ERR_NOMEM and self.active:
                    break                               # device fuller than the budget: wait
                self.queue.popleft()
                self._reject(req, str(e))
                continue
            self.queue.popleft()
            self.active.append([req, s, req._bytes, 0, self.clock()])
            self.used += req._bytes
            admitted += 1
        self.peak_used = max(self.peak_used, self.used)
        if admitted:
            self._dirty = True

    def _select(self) -> list:
        """The resident solves this round advances."""
        if self.policy == "all":
            return list(self.active)
        if self.policy == "hrrn":
            now = self.clock()
            # cell updates per second so far (the first rounds: a nominal rate)
            rate = self.cell_updates / now if self.cell_updates and now > 0 else 3e10

            def ratio(a):
                service = a[0].cells * (a[0].max_steps - a[3]) / rate
                return (now - a[0].arrival + service) / max(service, 1e-9)
            order = sorted(self.active, key=lambda a: (-ratio(a), a[4]))
        else:
            order = sorted(self.active,
                           key=lambda a: (a[0].cells * (a[0].max_steps - a[3]), a[4]))
        chosen, cells = [], 0
        for a in order:
            chosen.append(a)
            cells += a[0].cells
            if cells >= self.min_batch_cells:
                break
        if self.policy == "srpt_oldest":
            oldest = min(self.active, key=lambda a: a[4])
            if all(a is not oldest for a in chosen):
                chosen.append(oldest)
        return chosen

    def _rebuild(self, members):
        if self._batch.value:
            check(lib().lbm_batch_destroy(self._batch))
            self._batch = ctypes.c_void_p()
        self._batch_ids = tuple(a[0].id for a in members)
        if not members:
            return
        arr = (ctypes.c_void_p * len(members))(*[a[1].handle.value for a in members])
        check(lib().lbm_batch_create(arr, len(members), ctypes.byref(self._batch)))
        if self.tuning:
            t = self.tuning
            bt = _lib.lbm_batch_tuning(t["block"], t["cells_per_thread"], t["waves_per_cu"],
                                       (ctypes.c_int * 2)(*t["options"]), t.get("pairs", 0),
                                       t.get("group_mb", 0))
            check(lib().lbm_batch_set_tuning(self._batch, ctypes.byref(bt)))
        self._dirty = False

    def round(self) -> list:
        """Admit, run K steps of every member, check residuals, retire.
        Returns the results finished in this round."""
        now = self.clock()
        self._admit(now)
        done = []
        if self._cancel:
            keep = []
            for a in self.active:
                if a[0].id in self._cancel:
                    r = SolveResult(a[0].id, "cancelled", a[3], math.inf, a[0].arrival, a[4],
                                    self.clock())
                    self._drop(a)
                    self.results[r.id] = r
                    done.append(r)
                    self._dirty = True
                else:
                    keep.append(a)
            self.active = keep
            for req in [q for q in self.queue if q.id in self._cancel]:
                self.queue.remove(req)
                r = SolveResult(req.id, "cancelled", arrival=req.arrival, finished=self.clock())
                self.results[r.id] = r
                done.append(r)
            self._cancel.clear()
        if not self.active:
            return done
        members = self._select()
        if self._dirty or tuple(a[0].id for a in members) != self._batch_ids:
            self._rebuild(members)
            self._dirty = False
        st = _stream(self.stream)
        run = lib().lbm_batch_run if self.graph else lib().lbm_batch_launch
        check(run(self._batch, 0, self.K, st))
        res = (ctypes.c_double * len(members))()
        check(lib().lbm_batch_residual(self._batch, 0, res, st))
        self.rounds += 1
        finished = set()
        for i, a in enumerate(members):
            req, s = a[0], a[1]
            a[3] += self.K
            self.cell_updates += req.cells * self.K
            r = res[i]
            status = None
            # NaN: a non-finite state (any check); +inf: no snapshot yet (the
            # first check), later a vanishing field
            if math.isnan(r) or (math.isinf(r) and a[3] > self.K):
                status = "failed"
            elif r <= req.tol:
                status = "converged"
            elif a[3] >= req.max_steps:
                status = "step_cap"
            if status is None:
                continue
            out = _finish(req, s, status, a[3], r, self.stream)
            out.admitted = a[4]
            out.finished = self.clock()
            if status == "failed":
                out.message = "non-finite residual"
            finished.add(req.id)
            self.results[out.id] = out
            done.append(out)
        if finished:
            # the batch goes before its members' arrays
            check(lib().lbm_batch_destroy(self._batch))
            self._batch = ctypes.c_void_p()
            self._batch_ids = ()
            for a in self.active:
                if a[0].id in finished:
                    self._drop(a)
            self.active = [a for a in self.active if a[0].id not in finished]
            self._dirty = True
        return done

    def _drop(self, a):
        a[1].close()
        self.used -= a[2]

    def idle(self) -> bool:
        return not self.queue and not self.active

    def run(self, requests, progress=None) -> dict:
        """Open loop: requests are submitted as the serving clock passes their
        arrival times; returns {id: SolveResult} once all have finished."""
        pending = deque(sorted(requests, key=lambda r: r.arrival))
        self.t0 = time.perf_counter()
        last = 0.0
        while pending or not self.idle():
            now = self.clock()
            while pending and pending[0].arrival <= now:
                self.submit(pending.popleft())
            if not self.active and not [q for q in self.queue if q.arrival <= now]:
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
                continue
            self.round()
            if progress and self.clock() - last > 10.0:
                last = self.clock()
                progress(self)
        return self.results

    def close(self):
        if self._batch.value:
            lib().lbm_batch_destroy(self._batch)
            self._batch = ctypes.c_void_p()
        for a in self.active:
            self._drop(a)
        self.active = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def solve_sequential(requests, precision="f32", check_every=100, device=0, stream=None,
                     progress=None) -> dict:
    """Plain baseline: requests in arrival order, one at a time, each on its
    own solver with one launch per half-step; same residual test and results
    as the server. Waits for each arrival (open loop)."""
    if check_every % 2 or check_every <= 0:
        raise ValueError("check_every must be even (whole alpha + beta pairs)")
    L = lib()
    st = _stream(stream)
    results = {}
    t0 = time.perf_counter()
    last = 0.0
    for req in sorted(requests, key=lambda r: r.arrival):
        now = time.perf_counter() - t0
        if req.arrival > now:
            time.sleep(req.arrival - now)
        if req.params is None:
            try:
                req.params = cavity_parameters(req.size, req.u_lid, req.reynolds, precision)
            except LbmError as e:
                results[req.id] = SolveResult(req.id, "rejected", arrival=req.arrival,
                                              finished=req.arrival, message=str(e))
                continue
        admitted = time.perf_counter() - t0
        s = _new_solver(req, precision, device)
        steps, r, status = 0, math.inf, "step_cap"
        while steps < req.max_steps:
            for _ in range(check_every // 2):
                check(L.lbm_simulation_step_alpha(s.handle, None, None, st))
                check(L.lbm_simulation_step_beta(s.handle, None, None, st))
            steps += check_every
            r = s.residual(0, stream)
            if math.isnan(r) or (math.isinf(r) and steps > check_every):
                status = "failed"
                break
            if r <= req.tol:
                status = "converged"
                break
        out = _finish(req, s, status, steps, r, stream)
        s.close()
        out.admitted = admitted
        out.finished = time.perf_counter() - t0
        if status == "failed":
            out.message = "non-finite residual"
        results[req.id] = out
        if progress and time.perf_counter() - t0 - last > 10.0:
            last = time.perf_counter() - t0
            progress(len(results), len(requests))
    return results


def latency_summary(results) -> dict:
    """p50 / p99 / max latency (s) over finished requests, and solves/s over
    the span from the first arrival to the last finish."""
    ok = [r for r in results.values() if r.status in ("converged", "step_cap")]
    if not ok:
        return {"completed": 0}
    lat = np.array([r.latency for r in ok])
    span = max(r.finished for r in ok) - min(r.arrival for r in ok)
    return {
        "completed": len(ok),
        "p50_s": float(np.percentile(lat, 50)),
        "p99_s": float(np.percentile(lat, 99)),
        "max_s": float(lat.max()),
        "mean_s": float(lat.mean()),
        "span_s": float(span),
        "solves_per_s": len(ok) / span if span > 0 else math.inf,
    }
