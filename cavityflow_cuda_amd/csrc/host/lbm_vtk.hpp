// VTK output of one rank's interior (CLbmVisualizationVTK.cpp:69-205): the
// legacy ASCII structured grid with the subdomain's cell corners, then
// CELL_DATA flags, densities and velocities. Written through an ofstream with
// its default formatting, as the reference writes it, so the files match
// byte for byte.
#pragma once

#include <cstdint>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <typeinfo>

namespace lbm {

template <class T>
void write_vtk(const std::string& dir, int rank, int iteration, const int size[3],
               const int origin[3], const T length[3], const int32_t* flags, const T* rho,
               const T* u)
{
    std::stringstream name;
    name << dir << "/visualization_" << rank << "_" << iteration << ".vtk";
    std::ofstream file(name.str().c_str(), std::ios::out);
    if (!file)
        throw std::runtime_error("cannot open " + name.str());
    const char* tname = typeid(T) == typeid(float) ? "float" : "double";
    const long cells = (long)size[0] * size[1] * size[2];
    file << "# vtk DataFile Version 3.0\n";
    file << "Heterogenous (MPI/OpenMPI) and hybrid (CPU/GPU) LBM simulation on rank " << rank
         << "\n";
    file << "ASCII" << std::endl;
    file << "DATASET STRUCTURED_GRID\n";
    file << "DIMENSIONS " << (size[0] + 1) << " " << (size[1] + 1) << " " << (size[2] + 1) << "\n";
    file << "POINTS " << ((size[0] + 1) * (size[1] + 1) * (size[2] + 1)) << " " << tname << "\n";
    for (int k = 0; k < size[2] + 1; k++)
        for (int j = 0; j < size[1] + 1; j++)
            for (int i = 0; i < size[0] + 1; i++)
                file << ((T)(origin[0] + i) * length[0] / (T)size[0]) << " "
                     << ((T)(origin[1] + j) * length[1] / (T)size[1]) << " "
                     << ((T)(origin[2] + k) * length[2] / (T)size[2]) << "\n";
    file << "CELL_DATA " << cells << "\n";
    file << "SCALARS flags INT 1\n";
    file << "LOOKUP_TABLE default\n";
    for (long c = 0; c < cells; c++)
        file << flags[c] << "\n";
    file << "SCALARS densities " << tname << " 1\n";
    file << "LOOKUP_TABLE default\n";
    for (long c = 0; c < cells; c++)
        file << rho[c] << "\n";
    file << "VECTORS velocities " << tname << "\n";
    for (long c = 0; c < cells; c++)
        file << u[c] << " " << u[cells + c] << " " << u[2 * cells + c] << "\n";
}

}  // namespace lbm
