"""The native C++ driver (CLbmSolverHIP over the C ABI, CController-style
stepping and on-device face exchange) against the CPU oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

DRIVER = os.path.join(ROOT, "cavityflow_cuda_amd", "cavity_driver")


def test_driver_is_built_and_links():
    assert os.path.exists(DRIVER), "run __graft_entry__.build()"
    out = subprocess.run(["ldd", DRIVER], capture_output=True, text=True, check=True).stdout
    assert "libcavityflow_hip.so" in out and "not found" not in out


@pytest.mark.gpu
@pytest.mark.parametrize("prec,dtype", [("f32", np.float32), ("f64", np.float64)])
@pytest.mark.parametrize("split,fused_x", [((1, 1, 1), False), ((2, 1, 2), False),
                                           ((2, 2, 2), False), ((5, 1, 1), True)])
def test_driver_matches_oracle(oracle_cpu, tmp_path, prec, dtype, split, fused_x):
    """Undivided, split with the face exchange after each half-step, and
    split along x with the x faces exchanged inside the step kernels
    (--fused-x): the reference CPU kernels' cavity, bit for bit."""
    size, loops = (20, 16, 12), 41
    print("[case] native driver", prec, split, "fused x" if fused_x else "", flush=True)
    dump = str(tmp_path / "state.bin")
    p = subprocess.run([DRIVER, "--size", *map(str, size), "--split", *map(str, split),
                        "--u-lid", "0.05", "--re", "100", "--loops", str(loops),
                        "--precision", prec, "--dump", dump] + (["--fused-x"] if fused_x else []),
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "GLUPS" in p.stdout
    inv_tau = float([l for l in p.stdout.splitlines() if l.startswith("inv tau:")][0].split()[-1])
    raw = np.fromfile(dump, dtype=dtype)
    cells = int(np.prod(size))
    u, rho = raw[:3 * cells], raw[3 * cells:]
    o = oracle_cpu.OracleSolver(size, (1,) * 6, dtype=dtype, inv_tau=dtype(inv_tau),
                                lid_velocity=0.05)
    o.set_driven_cavity()
    for n in range(loops):
        o.beta() if n & 1 else o.alpha()
    assert u.tobytes() == o.getVelocities().tobytes()
    assert rho.tobytes() == o.getDensities().tobytes()
