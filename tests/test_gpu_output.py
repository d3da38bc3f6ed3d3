"""A run from an XML configuration writes the same visualization files (netCDF,
the reference's active writer, and VTK) from the native C++ driver and from
the Python controller (same kernels, same writer format), one file per rank
at step 0 and every `rate` steps."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def config_text(outdir):
    src = open(os.path.join(GOLDEN, "configs", "case0.xml")).read()
    src = src.replace("<domain-size><x>62</x><y>62</y><z>62</z></domain-size>",
                      "<domain-size><x>12</x><y>10</y><z>8</z></domain-size>")
    src = src.replace("<domian-length><x>62.0</x><y>62.0</y><z>62.0</z></domian-length>",
                      "<domian-length><x>1.2</x><y>1.0</y><z>0.8</z></domian-length>")
    src = src.replace("<subdomain-num><x>1</x>", "<subdomain-num><x>2</x>")
    src = src.replace("<loops>2</loops>", "<loops>6</loops>")
    src = src.replace("<visualization><do>0</do><rate>1</rate><output-dir>.</output-dir>",
                      f"<visualization><do>1</do><rate>3</rate><output-dir>{outdir}</output-dir>")
    assert src.count(outdir) == 1 and "<x>12</x>" in src
    return src


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_driver_and_controller_write_identical_vtk(tmp_path, hip, precision):
    from cavityflow_cuda_amd import build
    from cavityflow_cuda_amd.config import LbmConfiguration
    print("[case] vtk", precision, flush=True)
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    cpp_dir, py_dir = tmp_path / "cpp", tmp_path / "py"
    cpp_dir.mkdir()
    py_dir.mkdir()
    cfg_cpp = tmp_path / "cpp.xml"
    cfg_cpp.write_text(config_text(str(cpp_dir)))
    cfg_py = tmp_path / "py.xml"
    cfg_py.write_text(config_text(str(py_dir)))
    r = subprocess.run([exe, "--config", str(cfg_cpp), "--precision", precision,
                        "--vis-format", "vtk"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cfg = LbmConfiguration.from_xml(str(cfg_py), precision)
    c = hip.CavityController.from_config(cfg, vis_format="vtk")
    c.run(cfg.loops)
    c.close()
    names = sorted(os.listdir(py_dir))
    assert names == sorted(f"visualization_{r}_{i}.vtk" for r in (0, 1) for i in (0, 3, 6))
    assert sorted(os.listdir(cpp_dir)) == names
    for n in names:
        a = (cpp_dir / n).read_text()
        b = (py_dir / n).read_text()
        assert a == b, n
    # rank 1 sits at x = 6 cells of length 0.1: its first corner is at 0.6
    text = (py_dir / "visualization_1_6.vtk").read_text().splitlines()
    assert text[4] == "DIMENSIONS 7 11 9" and text[6].split()[0] == "0.6"
    assert "nan" not in (py_dir / "visualization_0_6.vtk").read_text()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_driver_and_controller_write_identical_netcdf(tmp_path, hip, precision):
    """netCDF is the default writer of both; each rank's file holds that
    rank's interior fields as its solver returns them, read back with an
    independent reader of the format."""
    from scipy.io import netcdf_file
    from cavityflow_cuda_amd import build
    from cavityflow_cuda_amd.config import LbmConfiguration
    print("[case] netcdf", precision, flush=True)
    T = np.float64 if precision == "f64" else np.float32
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    cpp_dir, py_dir = tmp_path / "cpp", tmp_path / "py"
    cpp_dir.mkdir()
    py_dir.mkdir()
    cfg_cpp = tmp_path / "cpp.xml"
    cfg_cpp.write_text(config_text(str(cpp_dir)))
    cfg_py = tmp_path / "py.xml"
    cfg_py.write_text(config_text(str(py_dir)))
    r = subprocess.run([exe, "--config", str(cfg_cpp), "--precision", precision],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cfg = LbmConfiguration.from_xml(str(cfg_py), precision)
    c = hip.CavityController.from_config(cfg)
    c.run(cfg.loops)
    c.sync()
    fields = {}
    for rank in (0, 1):
        s = c.solvers[(rank, 0, 0)]
        fields[rank] = (s.getFlags(), s.getDensities(), s.getVelocities())
    c.close()
    names = sorted(os.listdir(py_dir))
    assert names == sorted(f"visualization_{r}_{i}.nc" for r in (0, 1) for i in (0, 3, 6))
    assert sorted(os.listdir(cpp_dir)) == names
    for n in names:
        assert (cpp_dir / n).read_bytes() == (py_dir / n).read_bytes(), n
    for rank in (0, 1):
        f = netcdf_file(str(py_dir / f"visualization_{rank}_6.nc"), "r", mmap=False)
        flags, rho, u = fields[rank]
        assert f.dimensions == {"x": 6, "y": 10, "z": 8}
        assert f.variables["flags"].data.ravel().tolist() == np.asarray(flags).ravel().tolist()
        assert f.variables["densities"].data.astype(T).tobytes() == np.asarray(rho, T).tobytes()
        for k, name in enumerate(("velocities_x", "velocities_y", "velocities_z")):
            got = f.variables[name].data.astype(T).tobytes()
            assert got == np.asarray(u, T).reshape(3, -1)[k].tobytes(), (rank, name)
        # cells of length 0.1: centres 0.05, 0.15, ... from 0 in every rank's file
        assert abs(float(f.variables["x"].data[0]) - 0.05) < 1e-6
        assert abs(float(f.variables["x"].data[-1]) - 0.55) < 1e-6
        f.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_driver_and_controller_write_identical_global_netcdf(tmp_path, hip, precision):
    """--vis-format netcdf-global / vis_format="netcdf-global": one file per
    output step for the whole cavity (the reference's parallel build), equal
    to the gathered fields of both subdomains."""
    from scipy.io import netcdf_file
    from cavityflow_cuda_amd import build
    from cavityflow_cuda_amd.config import LbmConfiguration
    print("[case] netcdf-global", precision, flush=True)
    T = np.float64 if precision == "f64" else np.float32
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    cpp_dir, py_dir = tmp_path / "cpp", tmp_path / "py"
    cpp_dir.mkdir()
    py_dir.mkdir()
    cfg_cpp = tmp_path / "cpp.xml"
    cfg_cpp.write_text(config_text(str(cpp_dir)))
    cfg_py = tmp_path / "py.xml"
    cfg_py.write_text(config_text(str(py_dir)))
    r = subprocess.run([exe, "--config", str(cfg_cpp), "--precision", precision,
                        "--vis-format", "netcdf-global"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    cfg = LbmConfiguration.from_xml(str(cfg_py), precision)
    c = hip.CavityController.from_config(cfg, vis_format="netcdf-global")
    c.run(cfg.loops)
    c.sync()
    flags, rho, u = c.getFlags(), c.getDensities(), c.getVelocities()
    c.close()
    names = sorted(os.listdir(py_dir))
    assert names == sorted(f"visualization_{i}.nc" for i in (0, 3, 6))
    assert sorted(os.listdir(cpp_dir)) == names
    for n in names:
        assert (cpp_dir / n).read_bytes() == (py_dir / n).read_bytes(), n
    f = netcdf_file(str(py_dir / "visualization_6.nc"), "r", mmap=False)
    assert f.dimensions == {"x": 12, "y": 10, "z": 8}
    assert f.variables["flags"].data.tolist() == flags.tolist()
    assert f.variables["densities"].data.astype(T).tobytes() == np.asarray(rho, T).tobytes()
    for k, name in enumerate(("velocities_x", "velocities_y", "velocities_z")):
        assert f.variables[name].data.astype(T).tobytes() == np.asarray(u[k], T).tobytes()
    x = f.variables["x"].data
    assert abs(float(x[0]) - 0.05) < 1e-6 and abs(float(x[-1]) - 1.15) < 1e-6
    f.close()


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_driver_validation_mode(tmp_path, hip, precision):
    """validation do=1 (main.cpp:121-303): the split run is compared cell by
    cell with a one-subdomain run of the same configuration."""
    from cavityflow_cuda_amd import build
    print("[case] validation", precision, flush=True)
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    vdir = tmp_path / "val"
    vdir.mkdir()
    text = config_text(str(tmp_path / "unused"))
    text = text.replace("<visualization><do>1</do>", "<visualization><do>0</do>")
    text = text.replace("<validation><do>0</do><output-dir>.</output-dir>",
                        f"<validation><do>1</do><output-dir>{vdir}</output-dir>")
    text = text.replace("<subdomain-num><x>2</x><y>1</y><z>1</z>",
                        "<subdomain-num><x>2</x><y>2</y><z>1</z>")
    assert str(vdir) in text and "<y>2</y><z>1</z></subdomain-num>" in text
    cfg = tmp_path / "v.xml"
    cfg.write_text(text)
    r = subprocess.run([exe, "--config", str(cfg), "--precision", precision],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "validation:      0 inequalities" in r.stdout, r.stdout
    lines = (vdir / "validation_4.txt").read_text().splitlines()
    assert lines == ["number of inequalities: 0"]


def test_driver_benchmark_files(tmp_path, hip):
    """benchmark do=1 (CController.cpp:1451-1490): every rank writes the
    loops / time / iterations / GLUPS / GBytes/s lines, with bytes per cell
    = 19 x 2 populations + flag + 4 stored outputs."""
    from cavityflow_cuda_amd import build
    print("[case] benchmark files", flush=True)
    exe = os.path.join(os.path.dirname(build.__file__), "cavity_driver")
    bdir = tmp_path / "bench"
    bdir.mkdir()
    text = config_text(str(tmp_path / "vis"))
    (tmp_path / "vis").mkdir()
    text = text.replace("<benchmark><do>0</do><output-dir>.</output-dir>",
                        f"<benchmark><do>1</do><output-dir>{bdir}</output-dir>")
    assert str(bdir) in text
    cfg = tmp_path / "b.xml"
    cfg.write_text(text)
    r = subprocess.run([exe, "--config", str(cfg)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    report = [l for l in r.stdout.splitlines() if l.split(":")[0] in
              ("loops", "time", "iterations", "lattice updates", "bandwidth")]
    assert len(report) == 5 and report[0] == "loops:           6"
    for rank in (0, 1):
        assert (bdir / f"benchmark_2_{rank}.txt").read_text().splitlines() == report
    glups = float(report[3].split()[-1].replace("GLUPS", ""))
    gbs = float(report[4].split()[-1].replace("GBytes/s", ""))
    assert abs(gbs / glups - (19 * 2 * 4 + 4 + 4 * 4)) < 1e-3 * 172
