"""Benchmark of the D3Q19 alpha/beta A-A step on one MI355X.

Default workload: the lid-driven cavity at one rank's real subdomain of a
2x2x2 decomposition -- 512^3 interior cells (514^3 with the wall/halo layer,
10.3 GB of FP32 populations), Re 1000, lid velocity 0.1, FP32 -- with all six
face halos packed (CComm face shapes, 514 x 514 cells; the 5 populations per face
and phase that the receiving interior reads) and unpacked on the device after
every half-step. The y / z unpacks target a shadow subdomain of the same shape so
the cavity walls stay physical; for FP32 the x faces are packed and unpacked inside
the step kernels (lbm_set_step_pack / lbm_set_step_unpack), each face receiving its
own pack of the step pair before -- a wall reflected through the exchange path, which
leaves the cavity unchanged -- from one of two buffers while packing into the
other (also for F16S and BF16S; F16C and FP64 use the copy kernels for x too).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
                    [--precision f32|f64|f16s|bf16s|f16c] [--no-halo]
                    [--cpu-sample-steps 2] [--workload step|precision|batched|serving|ceiling]

Prints one JSON line. `value` is MLUPS = interior cells * K / t / 1e6 over the
K timed steps (populations resident in HBM, no host transfers in the timed
region). `roofline` reports the dominant kernel's algorithmic HBM rate:
2 * 19 * sizeof(population) bytes per interior cell per launch, divided by its
mean duration measured with HIP events on the launch stream. `cpu_baseline`
times the reference CPU kernels (compiled into oracle/_ref) or, where they
were not built, the oracle's C restatement, on a bounded sample of the same
workload on this host.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & %HBM roofline; served cavity solves/s + p99 latency vs plain"
BYTES_PER_POP = {"f32": 4, "f64": 8, "f16s": 2, "bf16s": 2, "f16c": 2}


class Hip:
    """Minimal HIP runtime calls (events, streams) through ctypes."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        for fn in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                   "hipEventElapsedTime", "hipEventDestroy", "hipStreamCreate",
                   "hipStreamSynchronize", "hipStreamDestroy", "hipDeviceSynchronize",
                   "hipMemcpyAsync", "hipGetDeviceCount", "hipSetDevice", "hipHostMalloc",
                   "hipHostFree"):
            getattr(self.L, fn).restype = ctypes.c_int

    def ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def stream(self):
        s = ctypes.c_void_p()
        self.ok(self.L.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        return s

    def event(self):
        e = ctypes.c_void_p()
        self.ok(self.L.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        return e

    def record(self, e, s):
        self.ok(self.L.hipEventRecord(e, s), "hipEventRecord")

    def elapsed_ms(self, a, b):
        ms = ctypes.c_float()
        self.ok(self.L.hipEventSynchronize(b), "hipEventSynchronize")
        self.ok(self.L.hipEventElapsedTime(ctypes.byref(ms), a, b), "hipEventElapsedTime")
        return ms.value

    def sync(self, s):
        self.ok(self.L.hipStreamSynchronize(s), "hipStreamSynchronize")

    def host_alloc(self, nbytes):
        """Pinned host memory."""
        p = ctypes.c_void_p()
        self.ok(self.L.hipHostMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes), 0),
                "hipHostMalloc")
        return p

    def host_free(self, p):
        self.ok(self.L.hipHostFree(p), "hipHostFree")


def measured_traffic(kernel, precision, size):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
    WRITE_SIZE profile of this workload (profiles/step_f32_512_hbm_traffic.md),
    or None when no profile covers it."""
    path = os.path.join(ROOT, "profiles", "step_f32_512_hbm_traffic.md")
    if precision != "f32" or size != 512 or not os.path.exists(path):
        return None
    import re
    m = re.search(rf"- {kernel}: ([0-9.]+) GB per launch", open(path).read())
    return int(float(m.group(1)) * 1e9) if m else None


def copy_bandwidth(hip, lib_, device, stream, gib=4.0, reps=10):
    """Device-to-device copy rate (read + write bytes / time) for context."""
    nbytes = int(gib * 2**30)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    from cavityflow_cuda_amd._lib import check
    check(lib_.lbm_device_alloc(device, nbytes, ctypes.byref(a)))
    check(lib_.lbm_device_alloc(device, nbytes, ctypes.byref(b)))
    e0, e1 = hip.event(), hip.event()
    hip.ok(hip.L.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, stream), "hipMemcpyAsync")
    hip.record(e0, stream)
    for _ in range(reps):
        hip.ok(hip.L.hipMemcpyAsync(b, a, ctypes.c_size_t(nbytes), 3, stream), "hipMemcpyAsync")
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    lib_.lbm_device_free(device, a)
    lib_.lbm_device_free(device, b)
    return 2 * nbytes * reps / (ms * 1e-3) / 1e9


def host_cpu():
    """CPU model, logical CPUs of the host and CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return dict(cpu_model=model, nproc=os.cpu_count(), usable_cpus=usable)


def cpu_baseline(size, params, sample_steps, threads_all):
    """Reference CPU kernels (oracle/_ref) on the same cavity, `sample_steps`
    whole-domain steps; returns (dict, dict-for-all-cores-oracle)."""
    from oracle import cpu, ref
    cells = int(np.prod(size))
    inv_tau, lid = params["inv_tau"], params["velocity_dimless"][0]
    out = {}
    if ref.available():
        s = ref.RefSolver(size, (1,) * 6, dtype=np.float32, inv_tau=inv_tau, lid_velocity=lid,
                          store_densities=False, store_velocities=False)
        nxh, nyh, nzh = s.size_h
        s.flags.reshape(nzh, nyh, nxh)[1:-1, nyh - 2, 1:-1] = 4
        kind, threads = "reference", 1
    else:
        s = cpu.OracleSolver(size, (1,) * 6, dtype=np.float32, inv_tau=inv_tau,
                             lid_velocity=lid, store_densities=False, store_velocities=False,
                             nthreads=1)
        s.set_driven_cavity()
        kind, threads = "port", 1
    t0 = time.perf_counter()
    for n in range(sample_steps):
        s.beta() if n & 1 else s.alpha()
    dt = time.perf_counter() - t0
    host = host_cpu()
    out["single"] = dict(value=round(cells * sample_steps / dt / 1e6, 3), unit="MLUPS",
                         cores=threads, kind=kind, **host,
                         sample=f"{sample_steps} whole-domain steps (alpha+beta) of the same "
                                f"{size[0]}x{size[1]}x{size[2]} FP32 cavity, "
                                f"{'reference CLbmAlphaCPU/CLbmBetaCPU' if kind == 'reference' else 'oracle C restatement'}"
                                f", {dt:.2f} s")
    del s
    if threads_all > 1:
        o = cpu.OracleSolver(size, (1,) * 6, dtype=np.float32, inv_tau=inv_tau, lid_velocity=lid,
                             store_densities=False, store_velocities=False,
                             nthreads=threads_all)
        o.set_driven_cavity()
        # a sample of about 10 s: whole steps (alpha + beta) until 10 s have
        # passed, at least the single-thread sample, at most 400 half-steps
        t0 = time.perf_counter()
        n = 0
        while n < 400 and (n < sample_steps or time.perf_counter() - t0 < 10.0):
            o.alpha()
            o.beta()
            n += 2
        dt = time.perf_counter() - t0
        sample_steps = n
        out["all"] = dict(value=round(cells * sample_steps / dt / 1e6, 3), unit="MLUPS",
                          cores=threads_all, kind="port", **host,
                          sample=f"{sample_steps} steps, oracle C restatement, OpenMP over z, "
                                 f"{dt:.2f} s")
    return out


def precision_workload(args, cf, L, hip, stream):
    """Shifted 16-bit storage against FP64 on the same cavity: 512^3 interior,
    lid 0.1, Re drawn from the seed in [400, 3200]; each run stops at a
    residual ||u^n - u^(n-100)|| / ||u^n|| <= 1e-6 or after --steps steps."""
    size = (args.size,) * 3
    rng = np.random.default_rng(args.seed)
    re = float(rng.uniform(400.0, 3200.0))
    cells = int(np.prod(size))
    runs = {}
    j64 = None
    for prec in ("f64", "f32", "f16s", "f16c", "bf16s"):
        params = cf.cavity_parameters(size, args.u_lid, re, prec)
        c = cf.CavityController(size, u_lid=args.u_lid, reynolds=re, precision=prec,
                                store_outputs=False, params=params, stream=stream.value)
        s = next(iter(c.solvers.values()))
        print(f"[precision] {prec} Re={re:.1f} tau={params['tau']:.5f}", file=sys.stderr,
              flush=True)
        e0, e1 = hip.event(), hip.event()
        hip.sync(stream)
        hip.record(e0, stream)
        w0 = time.perf_counter()
        last = w0
        steps, residual = 0, float("inf")
        while steps < args.steps:
            for _ in range(100):
                c.computeNextStep()
            steps += 100
            residual = s.residual(0, stream.value)   # even count: last step was beta
            if time.perf_counter() - last > 20:
                last = time.perf_counter()
                print(f"[precision] {prec} step {steps} residual {residual:.3e}",
                      file=sys.stderr, flush=True)
            if residual <= 1e-6:
                break
        hip.record(e1, stream)
        ms = hip.elapsed_ms(e0, e1)
        m = s.getMoments(0, stream=stream.value).reshape(4, -1)
        j = m[1:].astype(np.float64)
        rec = dict(steps=steps, residual=residual, mlups=round(cells * steps / (ms * 1e-3) / 1e6, 1),
                   tau=params["tau"])
        if prec == "f64":
            j64 = j
            norm = float(np.sqrt((j64 ** 2).sum()))
            rec["u_rms"] = float(np.sqrt((j64 ** 2).sum() / cells))
        else:
            d = j - j64
            rec["l2_rel"] = float(np.sqrt((d ** 2).sum()) / norm)
            rec["rms_abs"] = float(np.sqrt((d ** 2).sum() / cells))
            rec["linf_abs"] = float(np.abs(d).max())
            del d
        runs[prec] = rec
        print(f"[precision] {prec} {rec}", file=sys.stderr, flush=True)
        c.close()
        del m, j
    line = {
        "metric": "velocity L2 error vs FP64 (shifted 16-bit storage)",
        "value": runs["f16s"]["l2_rel"],
        "unit": "relative L2",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "f16s",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {size[0]}^3, lid {args.u_lid}, Re {re:.1f} "
                               f"(seed {args.seed}), to residual 1e-6 or {args.steps} steps",
                   "reynolds": re},
        "runs": runs,
        "mlups_ratio_f16s_over_f32": round(runs["f16s"]["mlups"] / runs["f32"]["mlups"], 3),
        "mlups_ratio_f16c_over_f32": round(runs["f16c"]["mlups"] / runs["f32"]["mlups"], 3),
    }
    print(json.dumps(line), flush=True)


def batched_workload(args, cf, L, hip, stream):
    """64 independent cavities (edges log-uniform in [32, 192], aspect <= 2,
    Re uniform in [100, 2000], lid uniform in [0.02, 0.1], seeded) stepped
    together by the batch launch with autotuned shapes and graph replay, in
    FP32 and shifted FP16, against plain sequential solves of the same
    cavities; every member is compared bit for bit with its solo run."""
    from cavityflow_cuda_amd.batch import CavityBatch, draw_cavities, solve_alone
    steps = args.steps
    runs = {}
    specs, redrawn = draw_cavities(args.cavities, args.seed)
    cells = sum(sp.cells for sp in specs)
    print(f"[batched] {len(specs)} cavities, {cells / 1e6:.2f} M interior cells, "
          f"{redrawn} draws redrawn, {steps} steps", file=sys.stderr, flush=True)
    for prec in ("f32", "f16s", "f16c"):
        for sp in specs:
            sp.params = cf.cavity_parameters(sp.size, sp.u_lid, sp.reynolds, prec)
        b = CavityBatch(specs, prec)
        # --no-graph: the trials too run as direct launches (a NULL stream is never captured)
        tuning = b.autotune(trial_steps=args.trial_steps,
                            stream=None if args.no_graph else stream.value)
        print(f"[batched] {prec} tuning {tuning}", file=sys.stderr, flush=True)
        b.run(100, stream=stream.value, graph=not args.no_graph)   # warm-up (graph capture)
        hip.sync(stream)
        b.reset()
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, stream)
        b.run(steps, stream=stream.value, graph=not args.no_graph)
        hip.record(e1, stream)
        ms = hip.elapsed_ms(e0, e1)
        batched = cells * steps / (ms * 1e-3) / 1e6
        print(f"[batched] {prec} batch {ms:.1f} ms, {batched:.0f} MLUPS", file=sys.stderr,
              flush=True)
        seq_s, same = 0.0, 0
        for i, sp in enumerate(specs):
            m, dt = solve_alone(sp, steps, prec, stream=stream.value)
            seq_s += dt
            same += int(m.tobytes() == b.moments(i).tobytes())
        seq = cells * steps / seq_s / 1e6
        print(f"[batched] {prec} sequential {seq_s:.2f} s, {seq:.0f} MLUPS, "
              f"{same}/{len(specs)} bitwise equal", file=sys.stderr, flush=True)
        pop = BYTES_PER_POP[prec]
        step_ms = ms / steps                  # one launch per member group per time step
        runs[prec] = dict(mlups=round(batched, 1), sequential_mlups=round(seq, 1),
                          speedup=round(batched / seq, 3), bitwise_equal=same,
                          batch_ms=round(ms, 2), sequential_s=round(seq_s, 3), tuning=tuning,
                          mean_step_ms=round(step_ms, 5),
                          hbm_gbs_alg=round(cells * 2 * 19 * pop / (step_ms * 1e-3) / 1e9, 1))
        b.close()
    r = runs["f32"]
    line = {
        "metric": "aggregate MLUPS, batched cavities",
        "value": r["mlups"], "unit": "MLUPS", "n_gpus": 1, "steps": steps, "warmup": 100,
        "ms_per_step": round(r["batch_ms"] / steps, 5), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{len(specs)} lid-driven cavities, edges log-uniform [32, 192], "
                               f"aspect <= 2, Re U[100, 2000], lid U[0.02, 0.1], seed {args.seed}, "
                               f"{steps} steps each", "interior_cells": cells,
                   "redrawn": redrawn, "graph_replay": not args.no_graph},
        "roofline": {"bound": "hbm", "achieved": r["hbm_gbs_alg"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(r["hbm_gbs_alg"] / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "runs": runs,
    }
    print(json.dumps(line), flush=True)


def serving_workload(args, cf, L, hip, stream):
    """Continuous-batching server against plain sequential solves on the same
    Poisson arrival trace (mean rate 80 % of the server's measured capacity):
    edges log-uniform in [48, 384] (aspect <= 2), Re uniform in [100, 3200],
    lid 0.1, each request to a residual of 1e-6 or the step cap, admission
    under an HBM budget."""
    from cavityflow_cuda_amd.batch import CavityBatch, CavitySpec
    from cavityflow_cuda_amd import serving as sv
    prec = args.precision
    reqs, redrawn = sv.draw_requests(args.requests, args.seed, max_steps=args.max_steps,
                                     precision=prec)
    work = np.array([r.cells * r.max_steps for r in reqs], dtype=np.float64)
    print(f"[serving] {len(reqs)} requests, mean {np.mean([r.cells for r in reqs]) / 1e6:.2f} M "
          f"cells, cap {args.max_steps} steps, {work.sum() / 1e12:.2f} T cell updates at the cap",
          file=sys.stderr, flush=True)
    # capacity: the batch rate on a sample of the requests, with autotuned shapes
    sample = [CavitySpec(r.size, r.u_lid, r.reynolds, r.params) for r in reqs[:24]]
    b = CavityBatch(sample, prec)
    # --no-graph: the trials too run as direct launches (a NULL stream is never captured)
    tuning = b.autotune(trial_steps=args.trial_steps,
                        stream=None if args.no_graph else stream.value)
    b.run(100, stream=stream.value, graph=not args.no_graph)
    hip.sync(stream)
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    b.run(300, stream=stream.value, graph=not args.no_graph)
    hip.record(e1, stream)
    cal = b.cells * 300 / (hip.elapsed_ms(e0, e1) * 1e-3)
    b.close()
    capacity = cal / work.mean()
    rate = args.load * capacity
    if args.load > 0:
        sv.set_poisson_arrivals(reqs, rate, args.seed)
    else:                       # closed loop: every request waiting at t = 0
        rate = math.inf
    print(f"[serving] calibration {cal / 1e6:.0f} MLUPS -> capacity {capacity:.3f} solves/s, "
          f"arrival rate {rate:.3f}/s, tuning {tuning}", file=sys.stderr, flush=True)

    def prog(srv):
        print(f"[serving] t={srv.clock():.0f}s done {len(srv.results)} active {len(srv.active)} "
              f"queued {len(srv.queue)} resident {srv.used / 1e9:.1f} GB", file=sys.stderr,
              flush=True)

    srv = sv.CavityServer(prec, hbm_budget=args.hbm_budget * 1e9, check_every=100,
                          stream=stream.value, tuning=tuning, policy=args.policy,
                          graph=not args.no_graph)
    served = srv.run(reqs, progress=prog)
    peak_gb = srv.peak_used / 1e9
    rounds = srv.rounds
    srv.close()
    s_sum = sv.latency_summary(served)
    s_lu = sum(reqs[i].cells * r.steps for i, r in served.items() if r.status != "rejected")
    s_mlups = s_lu / s_sum["span_s"] / 1e6
    print(f"[serving] served {s_sum}", file=sys.stderr, flush=True)

    def prog2(done, total):
        print(f"[serving] sequential done {done}/{total}", file=sys.stderr, flush=True)

    plain = sv.solve_sequential(reqs, prec, 100, stream=stream.value, progress=prog2)
    p_sum = sv.latency_summary(plain)
    p_lu = sum(reqs[i].cells * r.steps for i, r in plain.items() if r.status != "rejected")
    p_mlups = p_lu / p_sum["span_s"] / 1e6
    print(f"[serving] plain {prec} {p_sum}", file=sys.stderr, flush=True)
    # shifted storage is also compared with plain sequential solves in FP32, the
    # reference's precision (the same trace, its own CLbmSolver parameters)
    f32 = None
    if prec != "f32":
        reqs32, _ = sv.draw_requests(args.requests, args.seed, max_steps=args.max_steps)
        for r, r0 in zip(reqs32, reqs):
            r.arrival = r0.arrival
        plain32 = sv.solve_sequential(reqs32, "f32", 100, stream=stream.value, progress=prog2)
        f_sum = sv.latency_summary(plain32)
        f_lu = sum(reqs32[i].cells * r.steps for i, r in plain32.items() if r.status != "rejected")
        f32 = dict(f_sum, aggregate_mlups=round(f_lu / f_sum["span_s"] / 1e6, 1))
        f32["throughput_ratio_served_over_this"] = round(s_sum["solves_per_s"] / f_sum["solves_per_s"], 3)
        f32["p99_ratio_this_over_served"] = round(f_sum["p99_s"] / s_sum["p99_s"], 3)
        print(f"[serving] plain f32 {f_sum}", file=sys.stderr, flush=True)
    same = sum(int(served[i].steps == plain[i].steps and
                   served[i].totals.tobytes() == plain[i].totals.tobytes()) for i in served)
    statuses = {}
    for r in served.values():
        statuses[r.status] = statuses.get(r.status, 0) + 1
    line = {
        "metric": "served cavity solves/s (p50/p99 latency) vs plain sequential",
        "value": round(s_sum["solves_per_s"], 4), "unit": "solves/s", "n_gpus": 1,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": prec,
        "data": "synthetic",
        "config": {"workload": f"{len(reqs)} cavity solve requests, {prec} storage, " +
                               (f"Poisson arrivals at {args.load:.0%} of measured capacity"
                                if args.load > 0 else "all arriving at once") +
                               ", edges log-uniform [48, 384] "
                               f"(aspect <= 2), Re U[100, 3200], lid 0.1, residual 1e-6 or "
                               f"{args.max_steps} steps, HBM budget {args.hbm_budget:.0f} GB, "
                               f"seed {args.seed}, policy {args.policy}", "redrawn": redrawn,
                   "graph_replay": not args.no_graph,
                   "arrival_rate_per_s": round(rate, 4) if math.isfinite(rate) else None},
        "served": dict(s_sum, aggregate_mlups=round(s_mlups, 1), peak_resident_gb=round(peak_gb, 2),
                       rounds=rounds, statuses=statuses),
        "plain_sequential": dict(p_sum, aggregate_mlups=round(p_mlups, 1)),
        "plain_sequential_f32": f32,
        "throughput_ratio": round(s_sum["solves_per_s"] / p_sum["solves_per_s"], 3),
        "mlups_ratio": round(s_mlups / p_mlups, 3),
        "p99_ratio_plain_over_served": round(p_sum["p99_s"] / s_sum["p99_s"], 3),
        "bitwise_equal_results": f"{same}/{len(served)}",
        "calibration_mlups": round(cal / 1e6, 1),
    }
    print(json.dumps(line), flush=True)


def step_run(args, cf, L, hip, stream, precision, steps, warmup):
    """The step workload at one storage precision: warm-up, then `steps`
    timed steps with HIP events around every step kernel on the launch
    stream; returns the rates and per-kernel means."""
    from cavityflow_cuda_amd._lib import check
    device = 0
    size = (args.size,) * 3
    params = cf.cavity_parameters(size, args.u_lid, args.reynolds, precision)

    print(f"[bench] cavity {size} {precision} Re={args.reynolds} lid={args.u_lid} "
          f"tau={params['tau']:.6f}", file=sys.stderr, flush=True)
    c = cf.CavityController(size, u_lid=args.u_lid, reynolds=args.reynolds,
                            precision=precision, store_outputs=False, device=device,
                            params=params, stream=stream.value)
    solver = next(iter(c.solvers.values()))
    for k, opt in ((0, args.launch_alpha), (1, args.launch_beta)):
        if opt is not None:
            check(L.lbm_set_launch_options(solver.handle, k, opt))
    halo = not args.no_halo
    shadow = None
    bufs = []
    if halo:
        # shadow subdomain with the same halo planes: unpack target
        shadow = cf.LbmSolverHIP(size, (cf.GHOST_LAYER,) * 6, precision,
                                 store_densities=False, store_velocities=False, device=device)
        for d in range(6):
            p = ctypes.c_void_p()
            elem = 8 if precision == "f64" else 4
            check(L.lbm_device_alloc(device, solver.face_elements(d) * elem, ctypes.byref(p)))
            bufs.append(p)

    host_ex = halo and args.host_exchange
    hbufs = []
    if host_ex:
        elem = 8 if precision == "f64" else 4
        hbufs = [hip.host_alloc(solver.face_elements(d) * elem) for d in range(6)]
    fused = halo and not args.no_fused_pack and not host_ex
    if fused:
        # the step kernels pack the LEFT/RIGHT faces themselves
        check(L.lbm_set_step_pack(solver.handle, bufs[0], bufs[1]))
    # x faces exchanged inside the steps: each x face takes its received
    # populations from its own fused-pack buffer -- a wall reflected through
    # the exchange path (the pack stores what the wall's bounce-back returns,
    # so the cavity is unchanged), with the bytes and access pattern of a
    # neighbour's exchange -- and no x-face copy kernel runs
    # measured at 512^3: FP32 34.3 k MLUPS fused (alpha in 256-thread
    # workgroups at 3 waves per SIMD) vs 33.7 k with the copy kernels; F16S
    # 60.9 k (beta in 256-thread workgroups) vs 60.2 k; F16C 59.1 k vs 60.4 k;
    # FP64 17.1 k vs 17.6 k
    x_mode = args.x_exchange if args.x_exchange != "auto" else (
        "fused" if precision in ("f32", "f16s", "bf16s") else "kernels")
    x_in_steps = fused and x_mode == "fused"
    xbufs = []
    if x_in_steps:
        # two buffers per x face: each pair of steps packs into one and
        # unpacks from the other, as with a neighbour, whose steps write the
        # other phase's planes (other cache lines) of the buffer one reads
        elem = 8 if precision == "f64" else 4
        spare = []
        for d in (0, 1):
            p = ctypes.c_void_p()
            check(L.lbm_device_alloc(device, solver.face_elements(d) * elem, ctypes.byref(p)))
            spare.append(p)
        xbufs = [bufs[:2], spare]
        for d in (0, 1):
            for phase in (0, 1):
                check(L.lbm_pack_face(solver.handle, d, phase, xbufs[1][d], stream))
        check(L.lbm_set_step_unpack(solver.handle, xbufs[1][0], xbufs[1][1]))
    # as launched (defaults may depend on the step pack and unpack)
    launch_opts = []
    for k in (0, 1):
        v = ctypes.c_int()
        check(L.lbm_get_launch_options(solver.handle, k, ctypes.byref(v)))
        launch_opts.append(v.value)

    # x-face unpack: one paired call for 16-bit storage, one call per face for
    # FP32 / F64; with 5 planes per face the two measure the same at 512^3
    # (FP32 33,056 vs 33,084 MLUPS, F16S / F16C within 0.5 %).
    separate_unpack = (args.unpack == "separate" or
                       (args.unpack == "auto" and precision in ("f32", "f64")))

    def box3(box):
        return (ctypes.c_int * 3)(*box[0]), (ctypes.c_int * 3)(*box[1])

    def host_halo_exchange(phase):
        # the reference's calls on host buffers (CController.cpp:369-570,
        # 807-1008): after alpha the offset-1 plane goes into the neighbour's
        # halo plane, after beta the halo plane into its offset-1 plane, only
        # the inward directions; every call returns with its copy complete
        for axis in range(3):
            for d in (2 * axis, 2 * axis + 1):
                o, z = box3(solver.face_box(d, 0 if phase == 0 else 1))
                check(L.lbm_get_density_distributions(solver.handle, o, z, hbufs[d], 0, stream))
            for d in (2 * axis, 2 * axis + 1):
                r = d ^ 1
                o, z = box3(shadow.face_box(r, 1 if phase == 0 else 0))
                if phase == 0:
                    check(L.lbm_set_density_distributions(shadow.handle, o, z, hbufs[d], 0,
                                                          stream))
                else:
                    check(L.lbm_set_density_distributions_dir(shadow.handle, o, z, r, hbufs[d],
                                                              0, stream))

    yz_pairs = args.yz_copies == "pair"

    def halo_exchange(phase):
        if host_ex:
            return host_halo_exchange(phase)
        # reference order: x faces, then y, then z (CController.cpp:369-570)
        for axis in range(3):
            if x_in_steps and axis == 0:
                continue
            pair = yz_pairs if axis else not separate_unpack
            if not (fused and axis == 0):
                if axis and yz_pairs:
                    check(L.lbm_pack_face_pair(solver.handle, axis, phase, bufs[2 * axis],
                                               bufs[2 * axis + 1], stream))
                else:
                    for d in (2 * axis, 2 * axis + 1):
                        check(L.lbm_pack_face(solver.handle, d, phase, bufs[d], stream))
            if pair:
                # the low face receives what the high face sent, and vice versa
                check(L.lbm_unpack_face_pair(shadow.handle, axis, phase, bufs[2 * axis + 1],
                                             bufs[2 * axis], stream))
            else:
                for d in (2 * axis, 2 * axis + 1):
                    check(L.lbm_unpack_face(shadow.handle, d ^ 1, phase, bufs[d], stream))

    step_fns = (L.lbm_simulation_step_alpha, L.lbm_simulation_step_beta)
    counter = [0]

    def one_step(events=None):
        n = counter[0]
        phase = n & 1
        if x_in_steps:
            j = (n >> 1) & 1
            check(L.lbm_set_step_pack(solver.handle, xbufs[j][0], xbufs[j][1]))
            check(L.lbm_set_step_unpack(solver.handle, xbufs[1 - j][0], xbufs[1 - j][1]))
        if events is not None:
            hip.record(events[0], stream)
        check(step_fns[phase](solver.handle, None, None, stream))
        if events is not None:
            hip.record(events[1], stream)
        if halo:
            halo_exchange(phase)
        counter[0] += 1

    for _ in range(warmup):
        one_step()
    hip.sync(stream)

    # timed region: K steps, HIP events around every step kernel
    nev = steps
    ev = [(hip.event(), hip.event()) for _ in range(nev)]
    t_start, t_end = hip.event(), hip.event()
    hip.sync(stream)
    hip.record(t_start, stream)
    w0 = time.perf_counter()
    for i in range(steps):
        one_step(ev[i])
    hip.record(t_end, stream)
    hip.sync(stream)
    wall = time.perf_counter() - w0
    total_ms = hip.elapsed_ms(t_start, t_end)
    durs = np.array([hip.elapsed_ms(a, b) for a, b in ev])
    first_phase = warmup & 1
    alpha_ms = durs[(np.arange(nev) + first_phase) % 2 == 0]
    beta_ms = durs[(np.arange(nev) + first_phase) % 2 == 1]

    cells = int(np.prod(size))
    mlups = cells * steps / (total_ms * 1e-3) / 1e6
    bpp = BYTES_PER_POP[precision]
    alg_bytes = 2 * 19 * bpp * cells
    kern = {"alpha": float(alpha_ms.mean()), "beta": float(beta_ms.mean())}
    dom = max(kern, key=kern.get)
    achieved = alg_bytes / (kern[dom] * 1e-3) / 1e9
    traffic = measured_traffic(f"k_{dom}", precision, args.size)
    step_ms_total = float(durs.sum())
    halo_ms = total_ms - step_ms_total

    c.close()
    if shadow is not None:
        shadow.close()
    for p in bufs + (xbufs[1] if xbufs else []):
        L.lbm_device_free(device, p)
    for p in hbufs:
        hip.host_free(p)

    return dict(size=size, params=params, mlups=mlups, total_ms=total_ms, kern=kern, dom=dom,
                achieved=achieved, traffic=traffic, alg_bytes=alg_bytes, halo=halo, fused=fused,
                x_in_steps=x_in_steps,
                host_ex=host_ex, separate_unpack=separate_unpack, launch_opts=launch_opts,
                wall=wall, halo_ms=halo_ms, step_ms_total=step_ms_total, steps=steps)


PROBE_OPTION_NAMES = ((1, "NT"), (4, "WAVES4"), (8, "WAVES3"), (16, "WAVES2"), (256, "VEC2"),
                      (512, "VEC4"), (1024, "B512"), (2048, "B1024"), (4096, "SPAN2"),
                      (8192, "SPAN4"), (16384, "SPAN8"))


def probe_option_name(opt):
    return "+".join(n for bit, n in PROBE_OPTION_NAMES if opt & bit) or "none"


def ceiling_workload(args, cf, L, hip, stream):
    """Bandwidth ceilings of the step kernels' access patterns at the bench
    size and storage precision (`--precision`, FP32 by default): the library's
    pattern probes (lbm_probe_pattern: alpha's aligned 19-plane
    read-modify-write, beta's shifted one and its variants -- padded rows,
    x or y / z shifts only, and for FP32 every access line-aligned with the x
    shifts across lanes -- 19-plane read and write streams) under the launch
    options the steps use and under workgroup shapes of 256-1024 threads and
    1-8 cells per thread, a device-to-device copy, and
    the step kernels themselves (no exchange) for comparison. Each figure is
    bytes moved / mean launch time over `reps` launches timed with HIP events
    on the launch stream."""
    from cavityflow_cuda_amd._lib import check
    size = (args.size,) * 3
    prec = args.precision
    bpp = BYTES_PER_POP[prec]
    reps = 10
    s = cf.LbmSolverHIP(size, (cf.OBSTACLE,) * 6, prec, inv_tau=1.0 / 0.6536,
                        lid_velocity=args.u_lid, store_densities=False, store_velocities=False)
    s.setDrivenCavitySzenario()
    e0, e1 = hip.event(), hip.event()

    def timed(launch):
        launch()
        hip.sync(stream)
        hip.record(e0, stream)
        for _ in range(reps):
            launch()
        hip.record(e1, stream)
        return hip.elapsed_ms(e0, e1) / reps

    names = {0: "rmw_aligned", 1: "rmw_shifted", 2: "read", 3: "write",
             4: "rmw_shifted_pitched", 5: "rmw_shifted_x", 6: "rmw_shifted_yz",
             7: "rmw_lane_shifted"}
    vec = [256, 512, 256 | 16, 512 | 16, 256 | 8, 512 | 8, 256 | 4, 512 | 4, 512 | 1 | 16]
    # workgroup shapes: 256 / 512 / 1024 threads x 1 / 2 / 4 / 8 cells per
    # thread, uncapped or capped at 4 or 2 waves per SIMD (where one
    # workgroup fits), cached and nontemporal
    shapes = []
    for b in (0, 1024, 2048):
        for sp in (0, 4096, 8192, 16384):
            for cap in (0, 4, 16):
                if cap == 16 and b == 2048:
                    continue
                shapes += [b | sp | cap, b | sp | cap | 1]
    base = [0, 1, 16, 17, 8, 9, 4] + (vec if prec == "f32" else [])
    rmw = list(dict.fromkeys(base + shapes))
    # the lane-shifted pattern (FP32): 1-8 consecutive 64-cell chunks per wave
    lanes = [o for o in shapes if not (o & 2048 and o & 16384)]
    opt_sets = {0: rmw, 1: rmw, 2: [0, 1, 16], 3: [0, 1, 16], 4: [0, 16, 8, 4],
                5: [0, 16, 8, 4], 6: [0, 16, 8, 4], 7: lanes}
    modes = (2, 3, 0, 1, 4, 5, 6) + ((7,) if prec == "f32" else ())
    if args.patterns:
        want = set(args.patterns.split(",")) | {"rmw_aligned", "rmw_shifted"}
        unknown = want - set(names.values())
        if unknown:
            raise SystemExit(f"unknown patterns: {sorted(unknown)}")
        modes = tuple(m for m in modes if names[m] in want)
    if args.probe_options:
        chosen = [int(v) for v in args.probe_options.split(",")]
        # each pattern keeps the chosen values that its own sweep contains
        opt_sets = {m: [o for o in chosen if o in set(opt_sets[m])] for m in modes}
        for m in (0, 1):
            if not opt_sets[m]:
                raise SystemExit(f"--probe-options: none accepted by {names[m]}")
    probes = []
    for mode in modes:
        for opt in opt_sets[mode]:
            cells = ctypes.c_uint64()
            check(L.lbm_probe_pattern(s.handle, mode, opt, ctypes.byref(cells), stream))
            ms = timed(lambda: check(L.lbm_probe_pattern(s.handle, mode, opt, None, stream)))
            nbytes = cells.value * 19 * bpp * (1 if mode in (2, 3) else 2)
            probes.append(dict(pattern=names[mode], options=opt,
                               option_names=probe_option_name(opt), ms=round(ms, 4),
                               gbs=round(nbytes / (ms * 1e-3) / 1e9, 1)))
            print(f"[ceiling] {probes[-1]}", file=sys.stderr, flush=True)
    # the step kernels (whole domain, library default options, no exchange)
    s.reset()
    steps = {}
    for k, fn in ((0, L.lbm_simulation_step_alpha), (1, L.lbm_simulation_step_beta)):
        v = ctypes.c_int()
        check(L.lbm_get_launch_options(s.handle, k, ctypes.byref(v)))

        def pair():
            check(L.lbm_simulation_step_alpha(s.handle, None, None, stream))
            check(L.lbm_simulation_step_beta(s.handle, None, None, stream))
        pair()
        hip.sync(stream)
        ev = [(hip.event(), hip.event()) for _ in range(reps)]
        for a, b in ev:
            if k == 0:
                hip.record(a, stream)
                check(L.lbm_simulation_step_alpha(s.handle, None, None, stream))
                hip.record(b, stream)
                check(L.lbm_simulation_step_beta(s.handle, None, None, stream))
            else:
                check(L.lbm_simulation_step_alpha(s.handle, None, None, stream))
                hip.record(a, stream)
                check(L.lbm_simulation_step_beta(s.handle, None, None, stream))
                hip.record(b, stream)
        ms = float(np.mean([hip.elapsed_ms(a, b) for a, b in ev]))
        nbytes = int(np.prod(size)) * 2 * 19 * bpp
        steps["k_alpha" if k == 0 else "k_beta"] = dict(options=v.value, ms=round(ms, 4),
                                                      gbs=round(nbytes / (ms * 1e-3) / 1e9, 1))
    print(f"[ceiling] steps {steps}", file=sys.stderr, flush=True)
    s.close()
    copy = round(copy_bandwidth(hip, L, 0, stream), 1)
    best = {}
    for p in probes:
        if p["gbs"] > best.get(p["pattern"], {"gbs": 0})["gbs"]:
            best[p["pattern"]] = p
    line = {
        "metric": "HBM ceilings of the step kernels' access patterns",
        "value": best["rmw_shifted"]["gbs"], "unit": "GB/s", "n_gpus": 1,
        "higher_is_better": True, "dtype": prec, "data": "synthetic",
        "config": {"workload": f"{size[0]}^3 interior ({size[0] + 2}^3 with walls) {prec} "
                               f"population array, 19 planes of {bpp}-byte elements",
                   "reps": reps},
        "peak_gbs": HBM_PEAK_GBS,
        "device_copy_gbs": copy,
        "best": {k: dict(v, frac=round(v["gbs"] / HBM_PEAK_GBS, 4)) for k, v in best.items()},
        "probes": probes,
        "step_kernels": {k: dict(v, frac=round(v["gbs"] / HBM_PEAK_GBS, 4),
                                 of_pattern_ceiling=round(
                                     v["gbs"] / best["rmw_aligned" if k == "k_alpha"
                                                     else "rmw_shifted"]["gbs"], 4))
                         for k, v in steps.items()},
    }
    print(json.dumps(line), flush=True)


def split_workload(args, cf, L, hip, stream):
    """The cavity divided into subdomains on the one GPU (CavityController with
    `split`), faces exchanged on the device after every half-step: the serial
    order (whole subdomains, then the exchange) against the reference's
    boundary-first order with the exchange on a second stream
    (overlap=True) and, for a split along x only, the x faces exchanged inside
    the step kernels (fused_x=True); all give the same bits."""
    from cavityflow_cuda_amd import _hip
    size = (args.size,) * 3
    split = tuple(args.split)
    params = cf.cavity_parameters(size, args.u_lid, args.reynolds, args.precision)
    cells = int(np.prod(size))
    runs = {}
    variants = [("serial", {}), ("overlap_xslab2", dict(overlap=True, xslab=2))]
    if split[0] > 1:
        variants.append(("overlap_xslab32", dict(overlap=True, xslab=32)))
    if split[0] > 1 and split[1] == split[2] == 1:
        variants.append(("fused_x", dict(fused_x=True)))
    for key, kw in variants:
        main_stream = _hip.Stream(0)
        c = cf.CavityController(size, precision=args.precision, store_outputs=False,
                                split=split, params=params, stream=main_stream.handle, **kw)
        st = ctypes.c_void_p(main_stream.handle)
        for _ in range(args.warmup):
            c.computeNextStep()
        hip.sync(st)
        e0, e1 = hip.event(), hip.event()
        hip.record(e0, st)
        for _ in range(args.steps):
            c.computeNextStep()
        hip.record(e1, st)
        ms = hip.elapsed_ms(e0, e1)
        runs[key] = dict(mlups=round(cells * args.steps / (ms * 1e-3) / 1e6, 1),
                         ms_per_step=round(ms / args.steps, 4))
        print(f"[split] {key} {runs[key]}", file=sys.stderr, flush=True)
        c.close()
        main_stream.close()
    line = {
        "metric": "MLUPS, cavity split into subdomains on one GPU",
        "value": max(r["mlups"] for r in runs.values()), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "higher_is_better": True, "dtype": args.precision,
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {size[0]}^3, split {split[0]}x{split[1]}x"
                               f"{split[2]}, device-side face exchange every half-step",
                   "reynolds": args.reynolds, "u_lid": args.u_lid},
        "runs": runs,
        "overlap_over_serial": {k: round(v["mlups"] / runs["serial"]["mlups"], 4)
                                for k, v in runs.items() if k != "serial"},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--reynolds", type=float, default=1000.0)
    ap.add_argument("--u-lid", type=float, default=0.1)
    ap.add_argument("--precision", default="f32", choices=sorted(BYTES_PER_POP))
    ap.add_argument("--no-halo", action="store_true")
    ap.add_argument("--unpack", choices=("auto", "pair", "separate"), default="auto",
                    help="x-face unpack: both faces in one call (pair) or one call each; "
                         "auto = pair for 16-bit storage, separate otherwise")
    ap.add_argument("--yz-copies", choices=("pair", "separate"), default="pair",
                    help="y / z faces: both faces of an axis packed and unpacked in one "
                         "launch each (pair) or one launch per face")
    ap.add_argument("--host-exchange", action="store_true",
                    help="step: faces pass through pinned host buffers with the reference's "
                         "calls (getDensityDistributions / setDensityDistributions), "
                         "the PCIe-inclusive rate")
    ap.add_argument("--x-exchange", choices=("auto", "fused", "kernels"), default="auto",
                    help="x faces: packed and unpacked inside the step kernels (fused; each "
                         "face receives its own pack, a reflected wall) or unpacked by "
                         "lbm_unpack_face into the shadow subdomain (kernels); auto: fused "
                         "for FP32, F16S and BF16S, kernels for F16C and FP64 (measured "
                         "at 512^3)")
    ap.add_argument("--no-fused-pack", action="store_true",
                    help="pack the x faces with separate kernels instead of inside the steps")
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    # the GPU host gives one GPU's jobs a share of 16 CPUs (its OMP_NUM_THREADS)
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("OMP_NUM_THREADS", 0)) or
                    min(16, os.cpu_count() or 1))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-copy", action="store_true")
    ap.add_argument("--no-shifted", action="store_true",
                    help="step, f32: skip the F16S / F16C runs reported in `extra`")
    ap.add_argument("--launch-alpha", type=int, default=None,
                    help="LBM_LAUNCH_* bits for the alpha kernel (default: library default)")
    ap.add_argument("--launch-beta", type=int, default=None)
    ap.add_argument("--workload", default="step",
                    choices=("step", "precision", "batched", "serving", "ceiling", "split"))
    ap.add_argument("--split", type=int, nargs=3, default=(2, 2, 2),
                    help="split: subdomains per axis")
    ap.add_argument("--patterns", default=None,
                    help="ceiling: comma-separated probe patterns to run (rmw_aligned and "
                         "rmw_shifted always run; default all)")
    ap.add_argument("--probe-options", default=None,
                    help="ceiling: comma-separated lbm_probe_pattern option values to run "
                         "instead of the sweeps (each pattern takes the ones it accepts)")
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cavities", type=int, default=64, help="batched: number of cavities")
    ap.add_argument("--no-graph", action="store_true",
                    help="batched / serving: issue the batch launches directly instead of "
                         "replaying captured graphs")
    ap.add_argument("--trial-steps", type=int, default=8,
                    help="batched / serving: time steps per autotuner trial")
    ap.add_argument("--requests", type=int, default=200, help="serving: number of requests")
    ap.add_argument("--max-steps", type=int, default=2000,
                    help="serving: time-step cap per request")
    ap.add_argument("--load", type=float, default=0.8,
                    help="serving: arrival rate as a fraction of measured capacity "
                         "(0: all requests arrive at once)")
    ap.add_argument("--hbm-budget", type=float, default=200.0, help="serving: GB")
    ap.add_argument("--policy", default="hrrn", choices=("srpt", "srpt_oldest", "hrrn", "all"),
                    help="serving: which resident solves each round advances")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the multi-GPU path (RCCL face exchange "
              "between subdomains) is not part of this build; run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import cavityflow_cuda_amd as cf
    from cavityflow_cuda_amd import _lib
    from cavityflow_cuda_amd._lib import check

    hip = Hip()
    L = _lib.lib()
    device = 0
    stream = hip.stream()
    if args.workload == "precision":
        precision_workload(args, cf, L, hip, stream)
        return
    if args.workload == "batched":
        if args.steps == ap.get_default("steps"):
            args.steps = 5000
        batched_workload(args, cf, L, hip, stream)
        return
    if args.workload == "serving":
        serving_workload(args, cf, L, hip, stream)
        return
    if args.workload == "ceiling":
        ceiling_workload(args, cf, L, hip, stream)
        return
    if args.workload == "split":
        if args.steps == ap.get_default("steps"):
            args.steps = 200
        split_workload(args, cf, L, hip, stream)
        return
    run = step_run(args, cf, L, hip, stream, args.precision, args.steps, args.warmup)
    size, params, mlups, total_ms = run["size"], run["params"], run["mlups"], run["total_ms"]
    kern, dom, achieved, traffic = run["kern"], run["dom"], run["achieved"], run["traffic"]
    alg_bytes, halo, fused, host_ex = run["alg_bytes"], run["halo"], run["fused"], run["host_ex"]
    separate_unpack, launch_opts = run["separate_unpack"], run["launch_opts"]
    wall, halo_ms, step_ms_total = run["wall"], run["halo_ms"], run["step_ms_total"]

    # the shifted 16-bit formats on the same workload, for the ratio to FP32
    # (fewer steps: the rate is steady after a few hundred)
    shifted = {}
    if args.precision == "f32" and not args.no_shifted:
        for prec in ("f16s", "f16c"):
            r = step_run(args, cf, L, hip, stream, prec, min(args.steps, 1000), args.warmup)
            shifted[prec] = dict(mlups=round(r["mlups"], 1),
                                 ratio_to_f32=round(r["mlups"] / mlups, 3),
                                 steps=r["steps"],
                                 kernel_ms={k: round(v, 4) for k, v in r["kern"].items()},
                                 roofline_frac={k: round(r["alg_bytes"] / (v * 1e-3) / 1e9
                                                         / HBM_PEAK_GBS, 4)
                                                for k, v in r["kern"].items()},
                                 launch_options=r["launch_opts"])
            print(f"[bench] {prec}: {shifted[prec]}", file=sys.stderr, flush=True)

    copy_gbs = None
    if not args.skip_copy:
        copy_gbs = round(copy_bandwidth(hip, L, 0, stream), 1)

    cpu = {}
    if not args.skip_cpu and args.cpu_sample_steps > 0:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(size, params, args.cpu_sample_steps, args.cpu_threads)

    t_elem = 8 if args.precision == "f64" else 4
    line = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic",
        "config": {
            "workload": (f"lid-driven cavity, D3Q19 BGK A-A alpha/beta, {size[0]}^3 interior "
                         f"({size[0] + 2}^3 with walls), {args.precision} populations, "
                         f"Re {args.reynolds:g}, lid {args.u_lid:g}"
                         + ((", six-face exchange through pinned host buffers every half-step"
                             if host_ex else
                             (", six-face pack+unpack on the device every half-step (x faces "
                              "inside the step kernels, y / z faces by copy kernels)"
                              if run["x_in_steps"] else
                              ", six-face pack+unpack on the device every half-step"))
                            if halo else "")),
            "size": list(size),
            "reynolds": args.reynolds,
            "u_lid": args.u_lid,
            "tau": round(params["tau"], 6),
            "precision": args.precision,
            "halo_exchange": halo,
            "fused_x_pack": fused, "fused_x_unpack": run["x_in_steps"],
            "paired_unpack": halo and not separate_unpack and not host_ex and not run["x_in_steps"],
            "paired_yz_copies": halo and not host_ex and args.yz_copies == "pair",
            "host_exchange": host_ex,
            "parallelism": "single GPU (one rank's subdomain)",
            "launch_options": {"alpha": launch_opts[0], "beta": launch_opts[1]},
        },
        "roofline": {
            "bound": "hbm",
            "kernel": f"k_{dom}",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": ("rocprofv3 FETCH_SIZE + WRITE_SIZE per launch of this workload, "
                               "from the committed profile profiles/step_f32_512_hbm_traffic.md "
                               "(not measured in this run)") if traffic else None,
            # measured HBM bytes (rocprof, per launch) over this run's launch time
            "traffic_frac": (round(traffic / (kern[dom] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
                             if traffic else None),
            "bytes_per_launch": alg_bytes,
            "mean_ms": {k: round(v, 4) for k, v in kern.items()},
        },
        "cpu_baseline": cpu.get("single"),
        "cpu_baseline_all_cores": cpu.get("all"),
        "extra": {
            "wall_s": round(wall, 3),
            "halo_ms_per_step": round(halo_ms / args.steps, 4),
            "step_kernel_frac_of_time": round(step_ms_total / total_ms, 4),
            "device_copy_gbs": copy_gbs,
            "reference_glups": round(mlups / 1e3, 4),
            "reference_bandwidth_gbs": round(mlups / 1e3 * (19 * 2 * t_elem + 4), 1),
            "shifted_storage": shifted or None,
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
