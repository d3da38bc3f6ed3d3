"""Benchmark of the D3Q19 alpha/beta A-A step on one MI355X.

Default workload: the lid-driven cavity at one rank's real subdomain of a
2x2x2 decomposition -- 512^3 interior cells (514^3 with the wall/halo layer,
10.3 GB of FP32 populations), Re 1000, lid velocity 0.1, FP32 -- with all six
face halos packed (CComm shapes, 19 x 514 x 514) and unpacked on the device
after every half-step. The unpack targets a shadow subdomain of the same shape
so the cavity walls stay physical.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
                    [--precision f32|f64|f16s|bf16s] [--no-halo] [--cpu-sample-steps 2]

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
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS & %HBM roofline; served cavity solves/s + p99 latency vs plain"
BYTES_PER_POP = {"f32": 4, "f64": 8, "f16s": 2, "bf16s": 2}


class Hip:
    """Minimal HIP runtime calls (events, streams) through ctypes."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so.7")
        for fn in ("hipEventCreate", "hipEventRecord", "hipEventSynchronize",
                   "hipEventElapsedTime", "hipEventDestroy", "hipStreamCreate",
                   "hipStreamSynchronize", "hipStreamDestroy", "hipDeviceSynchronize",
                   "hipMemcpyAsync", "hipGetDeviceCount", "hipSetDevice"):
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
    out["single"] = dict(value=round(cells * sample_steps / dt / 1e6, 3), unit="MLUPS",
                         cores=threads, kind=kind,
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
        t0 = time.perf_counter()
        for n in range(sample_steps):
            o.beta() if n & 1 else o.alpha()
        dt = time.perf_counter() - t0
        out["all"] = dict(value=round(cells * sample_steps / dt / 1e6, 3), unit="MLUPS",
                          cores=threads_all, kind="port",
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
    for prec in ("f64", "f32", "f16s", "bf16s"):
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
    ap.add_argument("--no-fused-pack", action="store_true",
                    help="pack the x faces with separate kernels instead of inside the steps")
    ap.add_argument("--cpu-sample-steps", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1))
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-copy", action="store_true")
    ap.add_argument("--launch-alpha", type=int, default=None,
                    help="LBM_LAUNCH_* bits for the alpha kernel (default: library default)")
    ap.add_argument("--launch-beta", type=int, default=None)
    ap.add_argument("--workload", default="step", choices=("step", "precision"))
    ap.add_argument("--seed", type=int, default=20261202)
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
    size = (args.size,) * 3
    params = cf.cavity_parameters(size, args.u_lid, args.reynolds, args.precision)

    print(f"[bench] cavity {size} {args.precision} Re={args.reynolds} lid={args.u_lid} "
          f"tau={params['tau']:.6f}", file=sys.stderr, flush=True)
    c = cf.CavityController(size, u_lid=args.u_lid, reynolds=args.reynolds,
                            precision=args.precision, store_outputs=False, device=device,
                            params=params, stream=stream.value)
    solver = next(iter(c.solvers.values()))
    for k, opt in ((0, args.launch_alpha), (1, args.launch_beta)):
        if opt is not None:
            check(L.lbm_set_launch_options(solver.handle, k, opt))
    launch_opts = []
    for k in (0, 1):
        v = ctypes.c_int()
        check(L.lbm_get_launch_options(solver.handle, k, ctypes.byref(v)))
        launch_opts.append(v.value)
    halo = not args.no_halo
    shadow = None
    bufs = []
    if halo:
        # shadow subdomain with the same halo planes: unpack target
        shadow = cf.LbmSolverHIP(size, (cf.GHOST_LAYER,) * 6, args.precision,
                                 store_densities=False, store_velocities=False, device=device)
        for d in range(6):
            p = ctypes.c_void_p()
            elem = 8 if args.precision == "f64" else 4
            check(L.lbm_device_alloc(device, solver.face_elements(d) * elem, ctypes.byref(p)))
            bufs.append(p)

    fused = halo and not args.no_fused_pack
    if fused:
        # the step kernels pack the LEFT/RIGHT faces themselves
        check(L.lbm_set_step_pack(solver.handle, bufs[0], bufs[1]))

    def halo_exchange(phase):
        # reference order: x faces, then y, then z (CController.cpp:369-570)
        for axis in range(3):
            if not (fused and axis == 0):
                for d in (2 * axis, 2 * axis + 1):
                    check(L.lbm_pack_face(solver.handle, d, phase, bufs[d], stream))
            for d in (2 * axis, 2 * axis + 1):
                check(L.lbm_unpack_face(shadow.handle, d ^ 1, phase, bufs[d], stream))

    step_fns = (L.lbm_simulation_step_alpha, L.lbm_simulation_step_beta)
    counter = [0]

    def one_step(events=None):
        n = counter[0]
        phase = n & 1
        if events is not None:
            hip.record(events[0], stream)
        check(step_fns[phase](solver.handle, None, None, stream))
        if events is not None:
            hip.record(events[1], stream)
        if halo:
            halo_exchange(phase)
        counter[0] += 1

    for _ in range(args.warmup):
        one_step()
    hip.sync(stream)

    # timed region: K steps, HIP events around every step kernel
    nev = args.steps
    ev = [(hip.event(), hip.event()) for _ in range(nev)]
    t_start, t_end = hip.event(), hip.event()
    hip.sync(stream)
    hip.record(t_start, stream)
    w0 = time.perf_counter()
    for i in range(args.steps):
        one_step(ev[i])
    hip.record(t_end, stream)
    hip.sync(stream)
    wall = time.perf_counter() - w0
    total_ms = hip.elapsed_ms(t_start, t_end)
    durs = np.array([hip.elapsed_ms(a, b) for a, b in ev])
    first_phase = args.warmup & 1
    alpha_ms = durs[(np.arange(nev) + first_phase) % 2 == 0]
    beta_ms = durs[(np.arange(nev) + first_phase) % 2 == 1]

    cells = int(np.prod(size))
    mlups = cells * args.steps / (total_ms * 1e-3) / 1e6
    bpp = BYTES_PER_POP[args.precision]
    alg_bytes = 2 * 19 * bpp * cells
    kern = {"alpha": float(alpha_ms.mean()), "beta": float(beta_ms.mean())}
    dom = max(kern, key=kern.get)
    achieved = alg_bytes / (kern[dom] * 1e-3) / 1e9
    step_ms_total = float(durs.sum())
    halo_ms = total_ms - step_ms_total

    copy_gbs = None
    if not args.skip_copy:
        copy_gbs = round(copy_bandwidth(hip, L, device, stream), 1)

    cpu = {}
    if not args.skip_cpu and args.cpu_sample_steps > 0:
        print("[bench] CPU baseline sample", file=sys.stderr, flush=True)
        cpu = cpu_baseline(size, params, args.cpu_sample_steps, args.cpu_threads)

    c.close()
    if shadow is not None:
        shadow.close()
    for p in bufs:
        L.lbm_device_free(device, p)

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
                         + (", six-face pack+unpack on the device every half-step" if halo else "")),
            "size": list(size),
            "reynolds": args.reynolds,
            "u_lid": args.u_lid,
            "tau": round(params["tau"], 6),
            "precision": args.precision,
            "halo_exchange": halo,
            "fused_x_pack": fused,
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
            "traffic": measured_traffic(f"k_{dom}", args.precision, args.size),
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
        },
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
