"""netCDF output (the reference's active writer, CLbmVisualizationNetCDF.cpp):
no netCDF library is installed and the reference writer cannot be built, so
the files are pinned by an independent reader of the format (scipy.io's
netcdf_file) against the reference's layout — dimensions x, y, z; variables
x, y, z, flags, densities, velocities_x/_y/_z over (z, y, x) with their
units — and the C++ writer of the native driver must produce the same bytes
as the Python one."""
import io
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "cavityflow_cuda_amd", "csrc", "host")

UNITS = {"x": b"meters", "y": b"meters", "z": b"meters", "flags": b"flag",
         "densities": b"kg/m^3", "velocities_x": b"m/s", "velocities_y": b"m/s",
         "velocities_z": b"m/s"}


def fields(size, precision, seed):
    T = np.float64 if precision == "f64" else np.float32
    cells = int(np.prod(size))
    rng = np.random.default_rng(seed)
    flags = rng.choice(np.array([1, 2, 4, 8], np.int32), cells)
    rho = (1 + 0.01 * rng.standard_normal(cells)).astype(T)
    u = (0.05 * rng.standard_normal(3 * cells)).astype(T)
    u[::7] = -0.0                                    # signed zeros survive
    return flags, rho, u


def read(data):
    from scipy.io import netcdf_file
    return netcdf_file(io.BytesIO(data), "r", mmap=False)


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("offset64", [None, True])
def test_netcdf_layout_read_back_by_an_independent_reader(precision, offset64):
    from cavityflow_cuda_amd.output import netcdf_bytes
    T = np.float64 if precision == "f64" else np.float32
    size, length = (7, 5, 3), (0.7, 0.5, 0.3)
    flags, rho, u = fields(size, precision, 3)
    data = netcdf_bytes(size, length, flags, rho, u, precision, offset64=offset64)
    assert data[:4] == (b"CDF\x02" if offset64 else b"CDF\x01")
    f = read(data)
    assert f.dimensions == {"x": 7, "y": 5, "z": 3}
    assert list(f.variables) == list(UNITS)
    for name, v in f.variables.items():
        assert v.units == UNITS[name]
        assert v.dimensions == ((name,) if len(name) == 1 else ("z", "y", "x")), name
        assert v.data.dtype == (np.dtype(">i4") if name == "flags" else np.dtype(T).newbyteorder(">"))
    assert f.variables["flags"].data.ravel().tolist() == flags.tolist()
    for name, want in (("densities", rho), ("velocities_x", u[:105]),
                       ("velocities_y", u[105:210]), ("velocities_z", u[210:])):
        got = f.variables[name].data.astype(T).ravel()
        assert got.tobytes() == want.tobytes(), name
    # cell centres ((T)i + (T)0.5) * length / (T)size, evaluated in T
    for a, name in enumerate("xyz"):
        want = [T(T(T(i) + T(0.5)) * T(length[a])) / T(size[a]) for i in range(size[a])]
        assert f.variables[name].data.astype(T).tobytes() == np.array(want, T).tobytes()
    f.close()


def test_netcdf_header_bytes_follow_the_format():
    """The opening of the header spelled out from the classic-format grammar:
    magic, numrecs, the dimension list, the absent global attributes and the
    first variable entry up to its offset."""
    from cavityflow_cuda_amd.output import netcdf_bytes
    flags, rho, u = fields((2, 1, 1), "f32", 0)
    data = netcdf_bytes((2, 1, 1), (1.0, 1.0, 1.0), flags, rho, u)
    be = lambda *v: b"".join(int(x).to_bytes(4, "big") for x in v)
    want = (b"CDF\x01" + be(0) + be(10, 3) + be(1) + b"x\0\0\0" + be(2) + be(1) + b"y\0\0\0" +
            be(1) + be(1) + b"z\0\0\0" + be(1) + be(0, 0) + be(11, 8) + be(1) + b"x\0\0\0" +
            be(1, 0) + be(12, 1) + be(5) + b"unit" + b"s\0\0\0" + be(2) + be(6) + b"meters\0\0" +
            be(5, 8))
    assert data[:len(want)] == want
    begin_x = int.from_bytes(data[len(want):len(want) + 4], "big")
    # the variables follow one another from there: coordinates (2 + 1 + 1
    # floats), then five fields of 2 cells; the last one ends the file
    assert len(data) == begin_x + (2 + 1 + 1) * 4 + 5 * 2 * 4


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("nc") / "netcdf_harness")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", HOST, "-o", exe,
                    os.path.join(ROOT, "tests", "netcdf_harness.cpp")], check=True)
    return exe


@pytest.mark.parametrize("precision", ["f32", "f64"])
@pytest.mark.parametrize("variant", [0, 1, 2])
def test_native_writer_bytes_equal_python_writer(tmp_path, harness, precision, variant):
    from cavityflow_cuda_amd.output import NetCdfWriter, netcdf_bytes
    size, length = (6, 4, 5), (0.6, 0.4, 0.5)
    flags, rho, u = fields(size, precision, 7)
    raw = tmp_path / "in.bin"
    raw.write_bytes(flags.tobytes() + rho.tobytes() + u.tobytes())
    T = np.float64 if precision == "f64" else np.float32
    r = subprocess.run([harness, precision, str(tmp_path), "3", "120", *map(str, size),
                        *(repr(float(T(v))) for v in length), str(variant), str(raw)],
                       capture_output=True, text=True, check=True)
    path = r.stdout.strip()
    assert path == os.path.join(str(tmp_path), "visualization_3_120.nc")
    assert path == NetCdfWriter(3, 40, str(tmp_path), size, length, precision).path(120)
    want = netcdf_bytes(size, tuple(T(v) for v in length), flags, rho, u, precision,
                        offset64={0: None, 1: False, 2: True}[variant])
    assert open(path, "rb").read() == want


@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_native_global_file_bytes_equal_python_writer(tmp_path, harness, precision):
    """The parallel build's whole-domain file: coordinates over the whole
    extent from a subdomain's length and size."""
    from cavityflow_cuda_amd.output import NetCdfGlobalWriter, netcdf_bytes
    size, sub, sub_length = (12, 4, 10), (6, 4, 5), (0.3, 0.2, 0.25)
    flags, rho, u = fields(size, precision, 9)
    raw = tmp_path / "in.bin"
    raw.write_bytes(flags.tobytes() + rho.tobytes() + u.tobytes())
    T = np.float64 if precision == "f64" else np.float32
    r = subprocess.run([harness, precision, str(tmp_path), "-1", "40", *map(str, size),
                        *(repr(float(T(v))) for v in sub_length), "0", str(raw), *map(str, sub)],
                       capture_output=True, text=True, check=True)
    path = r.stdout.strip()
    assert path == os.path.join(str(tmp_path), "visualization_40.nc")
    assert path == NetCdfGlobalWriter(20, str(tmp_path), size, sub, sub_length, precision).path(40)
    data = open(path, "rb").read()
    assert data == netcdf_bytes(size, tuple(T(v) for v in sub_length), flags, rho, u, precision,
                                cell_size=sub)
    f = read(data)
    assert f.dimensions == {"x": 12, "y": 4, "z": 10}
    # cell length 0.3 / 6 = 0.05 on x: centres 0.025 ... 0.575 over both subdomains
    for a, name in enumerate("xyz"):
        want = [T(T(T(i) + T(0.5)) * T(sub_length[a])) / T(sub[a]) for i in range(size[a])]
        assert f.variables[name].data.astype(T).tobytes() == np.array(want, T).tobytes()
    f.close()
