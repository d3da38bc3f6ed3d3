"""The reference-side binding: include/CLbmSolverHIPRef.hpp is the
CLbmSolverHIP<T> : public CLbmSolver<T> subclass a maintainer adds to the
reference tree. Here it is compiled against the reference's own headers
(CLbmSolver.hpp, CConfiguration.hpp, CDomain.hpp, with CUDA's type headers as
the reference's common.h needs them) and linked with the reference's
CLbmSolver / CConfiguration / CDomain sources and libcavityflow_hip.so
(oracle/Makefile, target `binding`).

On the GPU the linked program reads the config-1 XML through the reference's
CConfiguration, lets CLbmSolver's constructor derive tau, sets the lid and
runs 1000 steps through the subclass's stream overloads: its velocities and
densities carry the SHA-256 of the compiled reference CPU kernels' run
(tests/golden/config1.json)."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

REF = "/root/reference/src"
BIN = os.path.join(ROOT, "oracle", "_ref", "ref_binding")


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_reference_binding_compiles_and_links():
    lib = os.path.join(ROOT, "cavityflow_cuda_amd", "libcavityflow_hip.so")
    if not os.path.exists(lib):
        pytest.skip("HIP library not built")
    out = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "binding"],
                         capture_output=True, text=True)
    assert out.returncode == 0, out.stderr
    assert "built" in out.stdout, out.stdout
    assert os.access(BIN, os.X_OK)
    # every C ABI symbol the subclass calls resolves in the library
    syms = subprocess.run(["nm", "-D", "--undefined-only", BIN], capture_output=True,
                          text=True).stdout
    used = {ln.split()[-1] for ln in syms.splitlines() if ln.split()[-1].startswith("lbm_")}
    assert {"lbm_create", "lbm_simulation_step_alpha", "lbm_get_velocities",
            "lbm_device_count"} <= used


@pytest.mark.gpu
def test_reference_binding_reproduces_config1(tmp_path):
    if not os.access(BIN, os.X_OK):
        pytest.skip("oracle/_ref/ref_binding not built (needs the reference sources)")
    print("[case] reference binding config1", flush=True)
    xml = open(os.path.join(GOLDEN, "configs", "case0.xml")).read()
    # validation on: CLbmSolver then stores rho and u (CLbmSolver.cpp:40-41)
    xml = xml.replace("<validation><do>0</do>", "<validation><do>1</do>")
    cfg = tmp_path / "config1.xml"
    cfg.write_text(xml)
    out = tmp_path / "out.bin"
    r = subprocess.run([BIN, str(cfg), "1000", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(GOLDEN, "config1.json")) as f:
        g = json.load(f)
    cells = 62 ** 3
    data = np.fromfile(out, dtype=np.float32)
    assert data.size == 4 * cells
    assert hashlib.sha256(data[:3 * cells].tobytes()).hexdigest() == g["sha256_u"]
    assert hashlib.sha256(data[3 * cells:].tobytes()).hexdigest() == g["sha256_rho"]
