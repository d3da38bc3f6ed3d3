"""VTK output of a subdomain's interior: the reference's legacy ASCII
structured-grid file (CLbmVisualizationVTK.cpp:69-205).

``visualization_<rank>_<iteration>.vtk`` holds the header, the grid points of
the subdomain's cell corners (origin + i) * length / size computed in the
arithmetic type, then CELL_DATA: flags (int), densities and velocities, in
the reference's order and number formatting (C++ stream defaults: ``%g`` with
6 significant digits). Fields come from the solver's stored outputs over the
interior box, as the reference's getters deliver them.
"""
from __future__ import annotations

import os

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
