"""VTK output of a subdomain's interior: the reference's legacy ASCII
structured-grid file (CLbmVisualizationVTK.cpp:69-205).

``visualization_<rank>_<iteration>.vtk`` holds the header, the grid points of
the subdomain's cell corners (origin + i) * length / size computed in the
arithmetic type, then CELL_DATA: flags (int), densities and velocities, in
the reference's order and number formatting (C++ stream defaults: ``%g`` with
6 significant digits). Fields come from the solver's stored outputs over the
interior box, as the reference's getters deliver them.

``visualization_<rank>_<iteration>.nc`` is the reference's active writer
(CLbmVisualizationNetCDF.cpp:92-228, chosen in CController.cpp:120-135) in the
netCDF classic format, written here without the netCDF library: dimensions
x, y, z; coordinate variables x(x), y(y), z(z) at the cell centres
(i + 0.5) * length / size in the arithmetic type; flags (int), densities and
velocities_x / _y / _z over (z, y, x); a "units" attribute on each. Big-endian
header, data right after it, variables in definition order, each padded to
4 bytes; the 64-bit-offset variant (version byte 2) when an offset or a
variable needs more than 31 bits. The parallel build's single file,
``visualization_<iteration>.nc``, holds the whole cavity instead.
"""
from __future__ import annotations

import os
import struct

import numpy as np


def _g(a) -> np.ndarray:
    return np.char.mod("%g", np.asarray(a, dtype=np.float64))


def vtk_text(rank, size, origin, length, flags, rho, u, precision="f32") -> str:
    """The file contents for interior fields flags / rho [z][y][x] and
    u [3][z][y][x] of a subdomain of `size` cells at `origin` (cells) with
    physical `length`."""
    T = np.float64 if precision == "f64" else np.float32
    tname = "double" if precision == "f64" else "float"
    nx, ny, nz = (int(v) for v in size)
    cells = nx * ny * nz
    out = ["# vtk DataFile Version 3.0\n",
           f"Heterogenous (MPI/OpenMPI) and hybrid (CPU/GPU) LBM simulation on rank {rank}\n",
           "ASCII\n",
           "DATASET STRUCTURED_GRID\n",
           f"DIMENSIONS {nx + 1} {ny + 1} {nz + 1}\n",
           f"POINTS {(nx + 1) * (ny + 1) * (nz + 1)} {tname}\n"]
    # corner coordinates in T: ((T)(origin + i) * length) / (T)size
    axes = []
    for a, n in enumerate((nx, ny, nz)):
        i = np.arange(n + 1)
        axes.append(_g((T(origin[a]) + i).astype(T) * T(length[a]) / T(n)))
    zz, yy, xx = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
    pts = np.char.add(np.char.add(np.char.add(np.char.add(xx.ravel(), " "), yy.ravel()), " "),
                      zz.ravel())
    out.append("\n".join(pts.tolist()) + "\n")
    out.append(f"CELL_DATA {cells}\n")
    out.append("SCALARS flags INT 1\nLOOKUP_TABLE default\n")
    out.append("\n".join(str(int(v)) for v in np.asarray(flags).ravel()) + "\n")
    out.append(f"SCALARS densities {tname} 1\nLOOKUP_TABLE default\n")
    out.append("\n".join(_g(np.asarray(rho, dtype=T).ravel()).tolist()) + "\n")
    out.append(f"VECTORS velocities {tname}\n")
    uu = np.asarray(u, dtype=T).reshape(3, cells)
    vs = np.char.add(np.char.add(np.char.add(np.char.add(_g(uu[0]), " "), _g(uu[1])), " "),
                     _g(uu[2]))
    out.append("\n".join(vs.tolist()) + "\n")
    return "".join(out)


class VtkWriter:
    """CLbmVisualizationVTK<T>: render(iteration) writes a file every `rate`
    iterations for one rank's solver."""

    def __init__(self, rank, rate, directory, size, origin, length, precision="f32"):
        self.rank, self.rate, self.directory = int(rank), int(rate), directory
        self.size, self.origin, self.length = tuple(size), tuple(origin), tuple(length)
        self.precision = precision

    def path(self, iteration) -> str:
        return os.path.join(self.directory, f"visualization_{self.rank}_{iteration}.vtk")

    def render(self, iteration, solver):
        if iteration % self.rate:
            return None
        flags = solver.getFlags()
        rho = solver.getDensities()
        u = solver.getVelocities()
        p = self.path(iteration)
        with open(p, "w") as f:
            f.write(vtk_text(self.rank, self.size, self.origin, self.length, flags, rho, u,
                             self.precision))
        return p


_NC_CHAR, _NC_INT, _NC_FLOAT, _NC_DOUBLE = 2, 4, 5, 6
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12


def _nc_text(s: str) -> bytes:
    b = s.encode()
    return struct.pack(">i", len(b)) + b + b"\0" * (-len(b) % 4)


def netcdf_bytes(size, length, flags, rho, u, precision="f32", offset64=None,
                 cell_size=None) -> bytes:
    """The .nc file contents for interior fields flags / rho [z][y][x] and
    u [3][z][y][x] over `size` cells (CLbmVisualizationNetCDF.cpp:92-155
    layout, 166-228 values). The coordinates are the cell centres
    ((T)i + 0.5) * length / (T)cell_size: `length` and `cell_size` are a
    subdomain's (cell_size defaults to `size`, one subdomain's file).
    offset64: None chooses the format variant by size; True / False force it."""
    T = np.dtype(np.float64 if precision == "f64" else np.float32)
    tt = _NC_DOUBLE if T == np.float64 else _NC_FLOAT
    nx, ny, nz = (int(v) for v in size)
    cn = tuple(int(v) for v in (cell_size if cell_size is not None else size))
    cells = nx * ny * nz
    coords = []
    for a, n in enumerate((nx, ny, nz)):     # ((T)i + (T)0.5) * length / (T)cell_size, in T
        i = np.arange(n).astype(T)
        coords.append((i + T.type(0.5)) * T.type(length[a]) / T.type(cn[a]))
    uu = np.asarray(u, dtype=T).reshape(3, cells)
    zyx = (2, 1, 0)
    vars_ = [("x", (0,), tt, "meters", coords[0].astype(T.newbyteorder(">"))),
             ("y", (1,), tt, "meters", coords[1].astype(T.newbyteorder(">"))),
             ("z", (2,), tt, "meters", coords[2].astype(T.newbyteorder(">"))),
             ("flags", zyx, _NC_INT, "flag", np.asarray(flags).reshape(cells).astype(">i4")),
             ("densities", zyx, tt, "kg/m^3",
              np.asarray(rho, dtype=T).reshape(cells).astype(T.newbyteorder(">"))),
             ("velocities_x", zyx, tt, "m/s", uu[0].astype(T.newbyteorder(">"))),
             ("velocities_y", zyx, tt, "m/s", uu[1].astype(T.newbyteorder(">"))),
             ("velocities_z", zyx, tt, "m/s", uu[2].astype(T.newbyteorder(">")))]
    vsize = [(v[4].nbytes + 3) & ~3 for v in vars_]
    if max(vsize) > 2**32 - 4:
        raise ValueError("netCDF: a variable of 4 GiB or more does not fit the classic format")
    last_begin = sum(vsize[:-1])
    if offset64 is None:
        offset64 = last_begin + 1024 > 2**31 - 1 or max(vsize) > 2**31 - 4
    head = [b"CDF" + bytes([2 if offset64 else 1]), struct.pack(">ii", 0, _NC_DIMENSION),
            struct.pack(">i", 3)]
    for name, n in (("x", nx), ("y", ny), ("z", nz)):
        head += [_nc_text(name), struct.pack(">i", n)]
    head += [struct.pack(">iii", 0, 0, _NC_VARIABLE), struct.pack(">i", len(vars_))]

    def entries(begin):
        out = []
        for (name, dims, t, units, _), vs in zip(vars_, vsize):
            out += [_nc_text(name), struct.pack(">i", len(dims)),
                    struct.pack(f">{len(dims)}i", *dims),
                    struct.pack(">ii", _NC_ATTRIBUTE, 1), _nc_text("units"),
                    struct.pack(">i", _NC_CHAR), _nc_text(units), struct.pack(">ii", t, vs),
                    struct.pack(">q" if offset64 else ">i", begin)]
            begin += vs
        return b"".join(out)

    fixed = b"".join(head)
    begin = len(fixed) + len(entries(0))
    parts = [fixed, entries(begin)]
    for (_, _, _, _, data), vs in zip(vars_, vsize):
        parts += [data.tobytes(), b"\0" * (vs - data.nbytes)]
    return b"".join(parts)


class NetCdfWriter:
    """CLbmVisualizationNetCDF<T> of the serial build: render(iteration)
    writes visualization_<rank>_<iteration>.nc every `rate` iterations with
    this rank's interior (the reference writes the fields at the subdomain's
    global origin into a file sized for the subdomain, which only rank 0's
    origin fits, and the coordinates from rank 0 only; here every rank's file
    holds its own fields and coordinates)."""

    def __init__(self, rank, rate, directory, size, length, precision="f32"):
        self.rank, self.rate, self.directory = int(rank), int(rate), directory
        self.size, self.length = tuple(size), tuple(length)
        self.precision = precision

    def path(self, iteration) -> str:
        return os.path.join(self.directory, f"visualization_{self.rank}_{iteration}.nc")

    def render(self, iteration, solver):
        if iteration % self.rate:
            return None
        p = self.path(iteration)
        with open(p, "wb") as f:
            f.write(netcdf_bytes(self.size, self.length, solver.getFlags(),
                                 solver.getDensities(), solver.getVelocities(), self.precision))
        return p


class NetCdfGlobalWriter:
    """CLbmVisualizationNetCDF<T> of the parallel build: render(iteration)
    writes visualization_<iteration>.nc every `rate` iterations with the whole
    cavity, each subdomain's fields at its global origin
    (CLbmVisualizationNetCDF.cpp:52, 107-111, 171-228), the coordinates from a
    subdomain's length and size."""

    def __init__(self, rate, directory, size, sub_size, sub_length, precision="f32"):
        self.rate, self.directory = int(rate), directory
        self.size, self.sub_size = tuple(size), tuple(sub_size)
        self.sub_length = tuple(sub_length)
        self.precision = precision

    def path(self, iteration) -> str:
        return os.path.join(self.directory, f"visualization_{iteration}.nc")

    def render(self, iteration, controller):
        if iteration % self.rate:
            return None
        p = self.path(iteration)
        with open(p, "wb") as f:
            f.write(netcdf_bytes(self.size, self.sub_length, controller.getFlags(),
                                 controller.getDensities(), controller.getVelocities(),
                                 self.precision, cell_size=self.sub_size))
        return p
