"""XML configuration and VTK output against the reference's own code:
CConfiguration fields and CLbmVisualizationVTK files written by the compiled
reference (tests/golden/configs, tests/golden/vtk, made by
tests/golden/make_golden.py), and the C library's atof / atoi."""
import ctypes
import json
import math
import os

import numpy as np
import pytest

from conftest import GOLDEN

CONFIGS = os.path.join(GOLDEN, "configs")
VTK = os.path.join(GOLDEN, "vtk")


@pytest.mark.parametrize("precision,dtype", [("f32", "float32"), ("f64", "float64")])
def test_configuration_fields_equal_reference(hip, precision, dtype):
    from cavityflow_cuda_amd.config import LbmConfiguration
    gold = json.load(open(os.path.join(CONFIGS, "configs.json")))
    for name, entry in gold.items():
        c = LbmConfiguration.from_xml(os.path.join(CONFIGS, name), precision)
        want = np.array(entry[dtype]["values"])
        got = np.array(c.fields())
        assert got.tobytes() == want.tobytes(), (name, got, want)     # -0.0 included
        assert c.visualization_dir == entry[dtype]["visualization_dir"]


@pytest.mark.parametrize("precision,dtype", [("f32", "float32"), ("f64", "float64")])
def test_solver_parameters_from_xml_equal_reference(hip, precision, dtype):
    from cavityflow_cuda_amd import LbmError
    from cavityflow_cuda_amd.config import LbmConfiguration
    params = json.load(open(os.path.join(GOLDEN, "params.json")))
    for i, entry in enumerate(params):
        c = LbmConfiguration.from_xml(os.path.join(CONFIGS, f"case{i}.xml"), precision)
        ref = entry[dtype]
        if ref["exit_code"] != 0:
            with pytest.raises(LbmError):
                c.solver_parameters()
            continue
        p = c.solver_parameters()
        got = [*p["velocity_dimless"], *p["acceleration_dimless"], p["viscosity_dimless"],
               p["tau"], p["inv_tau"], p["timestep"], p["viscosity"]]
        assert got == ref["values"][:11], (i, got, ref["values"])


def test_subdomains_follow_cmanager(hip):
    from cavityflow_cuda_amd.config import LbmConfiguration
    c = LbmConfiguration.from_xml(os.path.join(CONFIGS, "quirks.xml"), "f32")
    assert c.domain_size == (24, 12, 36) and c.subdomains == (2, 1, 3) and c.ranks == 6
    size, origin, length = c.subdomain(5)           # x fastest: (1, 0, 2)
    assert size == (12, 12, 12) and origin == (12, 0, 24)
    assert length == tuple(float(np.float32(v) / np.float32(n))
                           for v, n in zip(c.domain_length, c.subdomains))
    assert c.loops == -1 and c.visualization_rate == 7 and c.store_outputs


def test_check_parameters_rejects_like_the_reference(tmp_path, hip):
    from cavityflow_cuda_amd.config import LbmConfiguration
    src = open(os.path.join(CONFIGS, "case0.xml")).read()
    bad = {
        "non-cubic cells": src.replace("<domian-length><x>62.0</x>", "<domian-length><x>63.0</x>"),
        "size / subdomains": src.replace("<subdomain-num><x>1</x>", "<subdomain-num><x>4</x>"),
        "zero timestep": src.replace("<timestep>1.0</timestep>", "<timestep>0</timestep>"),
        "grid > 1024": src.replace("<alpha-grid-configuration><x>32</x><y>8</y>",
                                   "<alpha-grid-configuration><x>64</x><y>32</y>"),
    }
    for what, text in bad.items():
        assert text != src, what
        p = tmp_path / "bad.xml"
        p.write_text(text)
        with pytest.raises(ValueError):
            LbmConfiguration.from_xml(str(p))
    p = tmp_path / "missing.xml"
    p.write_text(src.replace("<viscosity>", "<visc>").replace("</viscosity>", "</visc>"))
    with pytest.raises(ValueError, match="physics/viscosity"):
        LbmConfiguration.from_xml(str(p))


def test_atof_atoi_match_the_c_library():
    from cavityflow_cuda_amd.config import atof, atoi
    libc = ctypes.CDLL(None)
    libc.atof.restype = ctypes.c_double
    libc.atof.argtypes = [ctypes.c_char_p]
    libc.atoi.restype = ctypes.c_int
    libc.atoi.argtypes = [ctypes.c_char_p]
    samples = ["0", " 12", "\t-7x", "+3.5", "24.9", "36abc", "-1e-3xyz", "1e", "1e+", ".5",
               "5.", "-.25e2", "abc", "", "  ", "0x1A", "0x1.8p1", "inf", "-Infinity", "nan",
               "1e400", "-0.0", "0.0123456789012345678", "7e-320", "+", "-", "1.5E+3q"]
    for s in samples:
        a = atof(s)
        b = libc.atof(s.encode())
        assert (math.isnan(a) and math.isnan(b)) or (a == b and math.copysign(1, a) ==
                                                      math.copysign(1, b)), (s, a, b)
        if "x" not in s.lower() and "inf" not in s.lower() and "nan" not in s.lower():
            assert atoi(s) == libc.atoi(s.encode()), s


def test_vtk_files_equal_reference(tmp_path, hip):
    from cavityflow_cuda_amd.output import vtk_text
    z = np.load(os.path.join(VTK, "inputs.npz"))
    for tag, prec in (("f32", "f32"), ("f64", "f64")):
        rank, it = (int(v) for v in z[f"{tag}_rank_iteration"])
        size = tuple(int(v) for v in z[f"{tag}_size"])
        text = vtk_text(rank, size, tuple(z[f"{tag}_origin"]), tuple(z[f"{tag}_length"]),
                        z[f"{tag}_flags"], z[f"{tag}_rho"], z[f"{tag}_u"], prec)
        want = open(os.path.join(VTK, f"visualization_{rank}_{it}.vtk")).read()
        assert text == want, tag


def _driver():
    import subprocess
    from cavityflow_cuda_amd import build
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    if not os.path.exists(exe):
        build.build_driver()
    return exe


@pytest.mark.parametrize("precision,dtype", [("f32", "float32"), ("f64", "float64")])
def test_native_driver_reads_configurations_like_the_reference(hip, precision, dtype):
    import subprocess
    exe = _driver()
    gold = json.load(open(os.path.join(CONFIGS, "configs.json")))
    params = json.load(open(os.path.join(GOLDEN, "params.json")))
    for name, entry in gold.items():
        p = subprocess.run([exe, "--config", os.path.join(CONFIGS, name), "--print-config",
                            "--precision", precision], capture_output=True, text=True)
        case = int(name[4]) if name.startswith("case") else None
        if case is not None and params[case][dtype]["exit_code"] != 0:
            assert p.returncode != 0 and "no valid timestep" in p.stderr
            continue
        assert p.returncode == 0, p.stderr
        lines = dict(l.split(": ", 1) for l in p.stdout.strip().splitlines())
        fields = np.array([float(v) for v in lines["fields"].split()])
        assert fields.tobytes() == np.array(entry[dtype]["values"]).tobytes(), name
        assert lines["visualization-dir"] == entry[dtype]["visualization_dir"]
        if case is not None:
            got = [float(v) for v in lines["params"].split()]
            assert got == params[case][dtype]["values"][:11], name


def test_native_driver_rejects_bad_configurations(tmp_path, hip):
    import subprocess
    exe = _driver()
    src = open(os.path.join(CONFIGS, "case0.xml")).read()
    for text in (src.replace("<domian-length><x>62.0</x>", "<domian-length><x>63.0</x>"),
                 src.replace("</physics>", ""), src.replace("<viscosity>", "<v>")):
        p = tmp_path / "bad.xml"
        p.write_text(text)
        r = subprocess.run([exe, "--config", str(p), "--print-config"], capture_output=True,
                           text=True)
        assert r.returncode != 0 and r.stderr.startswith("error:")


def test_native_driver_survives_malformed_xml(tmp_path, hip):
    """Truncated, cut and garbage-injected files end in an error status,
    never a crash (the reader was also run under ASan / UBSan)."""
    import random
    import subprocess
    exe = _driver()
    base = open(os.path.join(CONFIGS, "case0.xml")).read()
    rng = random.Random(7)
    for k in range(60):
        s = base
        if k % 3 == 0:
            s = s[:rng.randrange(len(s))]
        elif k % 3 == 1:
            i = rng.randrange(len(s))
            s = s[:i] + s[i + rng.randrange(1, 40):]
        else:
            i = rng.randrange(len(s))
            s = s[:i] + rng.choice(["<", ">", "</", "<!--", "&", "<a b='1'>", "/>"]) + s[i:]
        p = tmp_path / f"c{k}.xml"
        p.write_text(s)
        r = subprocess.run([exe, "--config", str(p), "--print-config"], capture_output=True,
                           text=True, timeout=30)
        assert r.returncode in (0, 1), (k, r.returncode, r.stderr[-300:])
        if r.returncode:
            assert r.stderr.startswith("error:"), r.stderr
