"""CPU-side checks of the C ABI: the library loads, exports every entry point
that include/lbm_abi.h declares, validates arguments before touching a GPU,
and its host-only parameter derivation matches the reference."""
import ctypes
import json
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

HEADER = os.path.join(ROOT, "include", "lbm_abi.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\**\s+\**(lbm_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lbm_create" in names and "lbm_simulation_step_beta" in names
    assert len(names) >= 25


def test_library_exports_every_declared_symbol(hip):
    from cavityflow_cuda_amd import _lib
    lib = _lib.lib()
    for name in declared_functions():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (lbm_\w+)", out))
    assert set(declared_functions()) <= exported
    assert set(_lib.EXPORTED) == set(declared_functions())


def test_create_rejects_bad_arguments_without_a_gpu(hip):
    from cavityflow_cuda_amd import _lib
    L = _lib.lib()
    h = ctypes.c_void_p()
    base = dict(size=(4, 4, 4), bc=(1,) * 6, precision=0)

    def params(**kw):
        d = dict(base, **kw)
        p = _lib.lbm_params()
        p.size[:] = d["size"]
        p.bc[:] = d["bc"]
        p.precision = d["precision"]
        p.inv_tau = 1.0
        return p

    for bad in (dict(size=(0, 4, 4)), dict(bc=(1, 1, 3, 1, 1, 1)), dict(precision=9),
                dict(size=(2000, 2000, 2000))):
        assert L.lbm_create(ctypes.byref(params(**bad)), ctypes.byref(h)) == -1
        assert L.lbm_last_error()
    assert L.lbm_create(None, ctypes.byref(h)) == -1
    assert L.lbm_destroy(None) == 0
    assert L.lbm_status_string(-2) == b"box outside the domain"


def test_face_exchanges_reject_bad_arguments_without_a_gpu(hip):
    """The buffer-less face exchanges check their arguments before any device
    work: NULL solvers, a negative count or a NULL list are argument errors,
    an empty list is a no-op."""
    from cavityflow_cuda_amd import _lib
    L = _lib.lib()
    assert L.lbm_exchange_face(None, 1, 0, None, None) == _lib.LBM_ERR_ARG
    arr = (_lib.lbm_face_transfer * 2)()
    arr[0].direction = arr[1].direction = 1
    assert L.lbm_exchange_faces(arr, 2, 0, None) == _lib.LBM_ERR_ARG
    assert b"invalid argument" in L.lbm_last_error()
    assert L.lbm_exchange_faces(arr, -1, 0, None) == _lib.LBM_ERR_ARG
    assert L.lbm_exchange_faces(None, 3, 1, None) == _lib.LBM_ERR_ARG
    assert L.lbm_exchange_faces(None, 0, 1, None) == _lib.LBM_OK


def test_compute_parameters_match_reference_golden(hip):
    """CLbmSolver constructor adaptation: bit-exact against the reference (host code)."""
    from cavityflow_cuda_amd import LbmError, compute_parameters
    with open(os.path.join(GOLDEN, "params.json")) as f:
        cases = json.load(f)
    for entry in cases:
        c = entry["case"]
        length = c.get("length") or c["size"]
        for name, prec, dtype in (("float32", "f32", np.float32), ("float64", "f64", np.float64)):
            ref = entry[name]
            kw = dict(length_x=length[0], size_x=c["size"][0], length_max=max(length),
                      velocity=c["velocity"], acceleration=c["acceleration"],
                      viscosity=c["viscosity"], timestep=c["timestep"],
                      max_velocity=c["max_vel"], max_acceleration=c["max_acc"], precision=prec)
            if ref["exit_code"] != 0:
                with pytest.raises(LbmError) as e:
                    compute_parameters(**kw)
                assert e.value.status == -5
                continue
            out = compute_parameters(**kw)
            v = np.array(ref["values"], dtype=dtype)
            assert (np.array(out["velocity_dimless"], dtype=dtype) == v[0:3]).all()
            assert (np.array(out["acceleration_dimless"], dtype=dtype) == v[3:6]).all()
            assert dtype(out["viscosity_dimless"]) == v[6]
            assert dtype(out["tau"]) == v[7]
            assert dtype(out["inv_tau"]) == v[8]
            assert dtype(out["timestep"]) == v[9]


def test_cavity_parameter_mapping(hip):
    """dx = dt = 1 mapping: u* = u_lid, tau = 0.5 (6 nu + 1), nu = u_lid N / Re."""
    from cavityflow_cuda_amd import cavity_parameters
    p = cavity_parameters((62, 62, 62), 0.05, 100.0)
    assert np.float32(p["velocity_dimless"][0]) == np.float32(0.05)
    assert abs(p["tau"] - 0.5 * (6 * 0.031 + 1)) < 1e-6
    p = cavity_parameters((512, 512, 512), 0.1, 1000.0)
    assert abs(p["tau"] - 0.6536) < 1e-6
