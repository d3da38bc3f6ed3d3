"""Write the golden vectors under tests/golden/ from the reference's own CPU
kernels (oracle/_ref/libref_lbm.so, compiled from /root/reference/src by
oracle/Makefile). Run in a container that holds the reference source tree:

    make -C oracle all && python tests/golden/make_golden.py

The reference has no test vectors of its own; these fixtures are its outputs
on seeded inputs:
  step_{f32,f64}.npz   random populations + mixed flags + gravity on a
                       non-cubic 6x5x7 interior: init, alpha, beta, and box steps
  flags_init.npz       CLbmInitCPU flag fields (its z-fastest enumeration) for
                       asymmetric boundary conditions on non-cubic domains
  cavity16.npz         16^3 lid-driven cavity (u_lid 0.05, Re 100), 200 steps
  config1.json         64^3-with-walls cavity (62^3 interior, u_lid 0.05,
                       Re 100), 1000 steps: SHA-256 of u / rho, plane sums and
                       the mid-z plane of u (config1_midplane.npz)
  params.json          CLbmSolver constructor outputs (tau, tauInv, u*, g*, dt)
                       for XML configurations that exercise its adaptation loop
"""
from __future__ import annotations

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import ref  # noqa: E402

SEED = 20261202
OBST, FLUID, INJ, GHOST = 1, 2, 4, 8


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def random_state(rng, size, dtype):
    n = int(np.prod([v + 2 for v in size]))
    dd = (rng.random(19 * n) * 0.08 + 0.01).astype(dtype)
    flags = rng.choice([OBST, FLUID, FLUID, FLUID, INJ, GHOST], size=n).astype(np.int32)
    return dd, flags


def step_fixture(dtype, name):
    rng = np.random.default_rng(SEED + (0 if dtype == np.float32 else 1))
    size = (6, 5, 7)
    kw = dict(dtype=dtype, inv_tau=1.6863406896591187, lid_velocity=0.05,
              gravitation=(0.0012, -0.0021, 0.0007))
    r = ref.RefSolver(size, (OBST,) * 6, **kw)
    out = dict(size=np.array(size), inv_tau=np.array(r.inv_tau), lid=np.array(r.lid),
               g=r.g.copy(), init_dd=r.dd.copy(), init_u=r.u.copy(), init_rho=r.rho.copy(),
               init_flags=r.flags.copy())
    dd0, flags = random_state(rng, size, dtype)
    r.dd[:] = dd0
    r.flags[:] = flags
    r.u[:] = 0
    r.rho[:] = 0
    out.update(dd0=dd0, flags=flags)
    r.alpha()
    out.update(alpha_dd=r.dd.copy(), alpha_u=r.u.copy(), alpha_rho=r.rho.copy())
    r.beta()
    out.update(beta_dd=r.dd.copy(), beta_u=r.u.copy(), beta_rho=r.rho.copy())
    # box steps on a sub-box that touches the halo
    origin, box = (0, 1, 2), (5, 6, 4)
    r.alpha(origin, box)
    r.beta(origin, box)
    out.update(box_origin=np.array(origin), box_size=np.array(box), box_dd=r.dd.copy(),
               box_u=r.u.copy(), box_rho=r.rho.copy())
    np.savez_compressed(os.path.join(HERE, name), **out)


def flags_fixture():
    out = {}
    cases = [((5, 3, 4), (OBST, GHOST, OBST, OBST, GHOST, OBST)),
             ((3, 4, 6), (GHOST, OBST, GHOST, OBST, OBST, GHOST)),
             ((4, 4, 4), (OBST,) * 6)]
    for i, (size, bc) in enumerate(cases):
        r = ref.RefSolver(size, bc, dtype=np.float32)
        out[f"size_{i}"] = np.array(size)
        out[f"bc_{i}"] = np.array(bc)
        out[f"flags_{i}"] = r.flags.astype(np.int8)
    np.savez_compressed(os.path.join(HERE, "flags_init.npz"), **out)


def cavity(size, u_lid, re, steps, dtype=np.float32):
    """Reference CPU run of the single-rank lid-driven cavity: CController
    whole-domain steps with the lid box of setDrivenCavitySzenario."""
    params = solver_params_for(size, u_lid, re, dtype)
    r = ref.RefSolver(size, (OBST,) * 6, dtype=dtype, inv_tau=params[8],
                      lid_velocity=params[0], gravitation=params[3:6])
    nxh, nyh, nzh = r.size_h
    flags = r.flags.reshape(nzh, nyh, nxh)
    flags[1:1 + size[2], nyh - 2, 1:1 + size[0]] = INJ
    for i in range(steps):
        if i & 1:
            r.beta()
        else:
            r.alpha()
    u = r.u.reshape(3, nzh, nyh, nxh)[:, 1:-1, 1:-1, 1:-1]
    rho = r.rho.reshape(nzh, nyh, nxh)[1:-1, 1:-1, 1:-1]
    return params, np.ascontiguousarray(u), np.ascontiguousarray(rho)


def xml_config(size, velocity, acceleration, viscosity, timestep, max_vel, max_acc,
               length=None):
    length = length or size
    tpl = """<?xml version="1.0"?>
<lbm-configuration>
  <physics>
    <velocity><x>{v[0]!r}</x><y>{v[1]!r}</y><z>{v[2]!r}</z></velocity>
    <acceleration><x>{a[0]!r}</x><y>{a[1]!r}</y><z>{a[2]!r}</z></acceleration>
    <viscosity>{nu!r}</viscosity>
    <max-velocity>{mv!r}</max-velocity>
    <max-acceleration>{ma!r}</max-acceleration>
  </physics>
  <grid>
    <domain-size><x>{s[0]}</x><y>{s[1]}</y><z>{s[2]}</z></domain-size>
    <domian-length><x>{l[0]!r}</x><y>{l[1]!r}</y><z>{l[2]!r}</z></domian-length>
    <subdomain-num><x>1</x><y>1</y><z>1</z></subdomain-num>
    <cpu-subdomain-ratio>1.0</cpu-subdomain-ratio>
  </grid>
  <simulation>
    <loops>2</loops>
    <timestep>{dt!r}</timestep>
    <benchmark><do>0</do><output-dir>.</output-dir></benchmark>
    <logging><do>0</do><output-dir>.</output-dir></logging>
    <validation><do>0</do><output-dir>.</output-dir></validation>
    <visualization><do>0</do><rate>1</rate><output-dir>.</output-dir></visualization>
  </simulation>
  <cpu><block-configuration>
    <alpha-block-configuration><x>32</x><y>32</y><z>32</z></alpha-block-configuration>
    <beta-block-configuration><x>32</x><y>32</y><z>32</z></beta-block-configuration>
  </block-configuration></cpu>
  <gpu><grid-configuration>
    <init-grid-configuration><x>32</x><y>16</y><z>1</z></init-grid-configuration>
    <alpha-grid-configuration><x>32</x><y>8</y><z>1</z></alpha-grid-configuration>
    <beta-grid-configuration><x>32</x><y>8</y><z>1</z></beta-grid-configuration>
  </grid-configuration></gpu>
</lbm-configuration>
"""
    return tpl.format(v=[float(x) for x in velocity], a=[float(x) for x in acceleration],
                      nu=float(viscosity), mv=float(max_vel), ma=float(max_acc), s=size,
                      l=[float(x) for x in length], dt=float(timestep))


def run_params(case, dtype):
    """CLbmSolver constructor via the reference, in a child process (the
    reference exits on infeasible inputs). Returns (exit code, 12 values)."""
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "cfg.xml")
        with open(path, "w") as f:
            f.write(xml_config(**case))
        code = ("import sys,json,numpy as np;sys.path.insert(0,%r);from oracle import ref;"
                "print(json.dumps(ref.solver_params(%r, np.%s).astype(float).tolist()))"
                % (ROOT, path, np.dtype(dtype).name))
        p = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True)
        vals = json.loads(p.stdout.strip().splitlines()[-1]) if p.returncode == 0 else None
        return p.returncode, vals


def solver_params_for(size, u_lid, re, dtype):
    n = max(size)
    case = dict(size=size, velocity=(u_lid, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
                viscosity=u_lid * n / re, timestep=1.0, max_vel=0.1, max_acc=0.1)
    rc, vals = run_params(case, dtype)
    assert rc == 0, "reference rejected the cavity parameters"
    return np.array(vals, dtype=dtype)


PARAM_CASES = [
    # config 1 mapping: no adaptation
    dict(size=(62, 62, 62), velocity=(0.05, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
         viscosity=0.05 * 62 / 100, timestep=1.0, max_vel=0.1, max_acc=0.1),
    # config 2 mapping
    dict(size=(512, 512, 512), velocity=(0.1, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
         viscosity=0.1 * 512 / 1000, timestep=1.0, max_vel=0.1, max_acc=0.1),
    # the reference default.xml physics (velocity-limited, gravity, auto viscosity)
    dict(size=(192, 192, 192), velocity=(1.0, 0.0, 0.0), acceleration=(0.0, -9.81, 0.0),
         viscosity=-1.0, timestep=0.01, max_vel=0.1, max_acc=0.1, length=(1.0, 1.0, 1.0)),
    # tau below 0.51 -> timestep re-scaled to TAU_DEFAULT
    dict(size=(32, 32, 32), velocity=(0.02, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
         viscosity=0.02 * 32 / 2000, timestep=1.0, max_vel=0.1, max_acc=0.1),
    # acceleration-limited
    dict(size=(48, 48, 48), velocity=(0.01, 0.0, 0.0), acceleration=(0.0, -0.5, 0.0),
         viscosity=0.05, timestep=1.0, max_vel=0.1, max_acc=0.1),
    # infeasible: velocity limit forces dt down, tau limit wants it up -> exit
    dict(size=(16, 16, 16), velocity=(1.0, 0.0, 0.0), acceleration=(0.0, 0.0, 0.0),
         viscosity=1e-6, timestep=1.0, max_vel=0.1, max_acc=0.1),
]


def params_fixture():
    out = []
    for case in PARAM_CASES:
        entry = dict(case=case)
        for dtype in (np.float32, np.float64):
            rc, vals = run_params(case, dtype)
            entry[np.dtype(dtype).name] = dict(exit_code=rc, values=vals)
        out.append(entry)
    with open(os.path.join(HERE, "params.json"), "w") as f:
        json.dump(out, f, indent=1)


QUIRKS_XML = """<?xml version="1.0"?>
<lbm-configuration>
  <physics>
    <velocity><x> 0.05 </x><y>0</y><z>-0.0</z></velocity>
    <acceleration><x>0.0</x><y>-1e-3xyz</y><z>0</z></acceleration>
    <viscosity>0.0123456789012345678</viscosity>
    <max-velocity>0.1</max-velocity>
    <max-acceleration>0.1</max-acceleration>
  </physics>
  <grid>
    <domain-size><x>24.9</x><y> 12</y><z>36abc</z></domain-size>
    <domian-length><x>2.4</x><y>1.2</y><z>3.6</z></domian-length>
    <subdomain-num><x>2</x><y>1</y><z>3</z></subdomain-num>
    <cpu-subdomain-ratio>0.25</cpu-subdomain-ratio>
  </grid>
  <simulation>
    <loops>-1</loops>
    <timestep>0.001</timestep>
    <benchmark><do>1</do><output-dir>bench_out</output-dir></benchmark>
    <logging><do>0</do><output-dir>log_out</output-dir></logging>
    <validation><do>1</do><output-dir>val_out</output-dir></validation>
    <visualization><do>1</do><rate>7</rate><output-dir>vis out/dir</output-dir></visualization>
  </simulation>
  <cpu><block-configuration>
    <alpha-block-configuration><x>16</x><y>8</y><z>4</z></alpha-block-configuration>
    <beta-block-configuration><x>8</x><y>4</y><z>2</z></beta-block-configuration>
  </block-configuration></cpu>
  <gpu><grid-configuration>
    <init-grid-configuration><x>64</x><y>4</y><z>1</z></init-grid-configuration>
    <alpha-grid-configuration><x>128</x><y>2</y><z>1</z></alpha-grid-configuration>
    <beta-grid-configuration><x>256</x><y>1</y><z>1</z></beta-grid-configuration>
  </grid-configuration></gpu>
</lbm-configuration>
"""


def configs_fixture():
    """The XML files behind params.json plus one exercising atoi / atof
    prefixes, whitespace and the MPI subdomain count, with every field
    CConfiguration<T> reads from each (configs/configs.json)."""
    d = os.path.join(HERE, "configs")
    os.makedirs(d, exist_ok=True)
    files = []
    for i, case in enumerate(PARAM_CASES):
        name = f"case{i}.xml"
        with open(os.path.join(d, name), "w") as f:
            f.write(xml_config(**case))
        files.append(name)
    with open(os.path.join(d, "quirks.xml"), "w") as f:
        f.write(QUIRKS_XML)
    files.append("quirks.xml")
    # the configuration files the reference ships (data, copied as fixtures)
    shipped = os.path.join(ref.REF_ROOT, "configurations")
    for name in sorted(os.listdir(shipped)):
        if name.endswith(".xml"):
            with open(os.path.join(shipped, name)) as f:
                text = f.read()
            with open(os.path.join(d, "reference_" + name), "w") as f:
                f.write(text)
            files.append("reference_" + name)
    out = {}
    for name in files:
        entry = {}
        for dtype in (np.float32, np.float64):
            vals, vis = ref.config(os.path.join(d, name), dtype)
            entry[np.dtype(dtype).name] = dict(values=vals[:41].tolist(), visualization_dir=vis)
        out[name] = entry
    with open(os.path.join(d, "configs.json"), "w") as f:
        json.dump(out, f, indent=1)


def vtk_fixture():
    """CLbmVisualizationVTK output for given interior fields: FP32 on a first
    subdomain, FP64 on an offset subdomain with non-unit lengths; values span
    signs, zeros and exponents so the number formatting is pinned."""
    d = os.path.join(HERE, "vtk")
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(SEED + 7)
    cases = [("f32", np.float32, (3, 4, 2), (0, 0, 0), (3.0, 4.0, 2.0), 0, 0),
             ("f64", np.float64, (5, 2, 3), (5, 0, 2), (1.0, 0.4, 0.6), 3, 120)]
    arrays = {}
    for tag, dtype, size, origin, length, rank, it in cases:
        n = int(np.prod(size))
        flags = rng.choice(np.array([1, 2, 4, 8], np.int32), n)
        rho = (rng.standard_normal(n) * 10.0 ** rng.integers(-12, 12, n)).astype(dtype)
        u = (rng.standard_normal(3 * n) * 10.0 ** rng.integers(-20, 3, 3 * n)).astype(dtype)
        rho[:3] = [0.0, -0.0, 1.0]
        u[:4] = [0.0, -0.0, 123456789.0, 1e-5]
        ref.vtk(os.path.join(HERE, "configs", "case0.xml"), d, rank, it, size, origin, length,
                flags, rho, u, dtype)
        arrays.update({f"{tag}_flags": flags, f"{tag}_rho": rho, f"{tag}_u": u,
                       f"{tag}_size": np.array(size), f"{tag}_origin": np.array(origin),
                       f"{tag}_length": np.array(length, dtype=dtype),
                       f"{tag}_rank_iteration": np.array([rank, it])})
    np.savez_compressed(os.path.join(d, "inputs.npz"), **arrays)


def main():
    if not ref.available():
        raise SystemExit("oracle/_ref/libref_lbm.so missing: run `make -C oracle ref` first")
    if len(sys.argv) > 1:             # only the named fixtures, e.g. `configs vtk`
        for name in sys.argv[1:]:
            globals()[f"{name}_fixture"]()
        return
    configs_fixture()
    vtk_fixture()
    step_fixture(np.float32, "step_f32.npz")
    step_fixture(np.float64, "step_f64.npz")
    flags_fixture()
    params_fixture()
    params, u, rho = cavity((16, 16, 16), 0.05, 100.0, 200)
    np.savez_compressed(os.path.join(HERE, "cavity16.npz"), u=u, rho=rho,
                        inv_tau=np.array(params[8]))
    params, u, rho = cavity((62, 62, 62), 0.05, 100.0, 1000)
    mid = u[:, 31].copy()
    np.savez_compressed(os.path.join(HERE, "config1_midplane.npz"), u_mid_z31=mid)
    with open(os.path.join(HERE, "config1.json"), "w") as f:
        json.dump(dict(size=[62, 62, 62], u_lid=0.05, reynolds=100.0, steps=1000,
                       inv_tau=float(params[8]), tau=float(params[7]),
                       sha256_u=sha(u), sha256_rho=sha(rho),
                       u_plane_sums=u.astype(np.float64).sum(axis=(2, 3)).tolist(),
                       rho_sum=float(rho.astype(np.float64).sum()),
                       max_abs_u=float(np.abs(u).max())), f, indent=1)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
