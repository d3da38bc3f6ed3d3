// Writes one netCDF file through lbm::write_netcdf from raw little-endian
// inputs, for the byte comparison with the Python writer:
//   netcdf_harness f32|f64 DIR RANK ITER NX NY NZ LX LY LZ VARIANT INPUT [SX SY SZ]
// INPUT holds flags (int32) [z][y][x], rho [z][y][x] and u [3][z][y][x] in the
// arithmetic type, one after the other. With SX SY SZ (a subdomain's size; LX
// LY LZ are then its length) the whole-domain file of write_netcdf_global.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "lbm_netcdf.hpp"

template <class T>
static int run(char** a, bool global)
{
    const std::string dir = a[2];
    const int rank = std::atoi(a[3]), iter = std::atoi(a[4]);
    const int size[3] = {std::atoi(a[5]), std::atoi(a[6]), std::atoi(a[7])};
    const T length[3] = {(T)std::atof(a[8]), (T)std::atof(a[9]), (T)std::atof(a[10])};
    const int variant = std::atoi(a[11]);
    const size_t cells = (size_t)size[0] * size[1] * size[2];
    std::vector<int32_t> flags(cells);
    std::vector<T> rho(cells), u(3 * cells);
    std::ifstream in(a[12], std::ios::binary);
    in.read((char*)flags.data(), cells * 4);
    in.read((char*)rho.data(), cells * sizeof(T));
    in.read((char*)u.data(), 3 * cells * sizeof(T));
    if (!in)
        return 2;
    if (global) {
        const int sub[3] = {std::atoi(a[13]), std::atoi(a[14]), std::atoi(a[15])};
        std::printf("%s\n", lbm::write_netcdf_global<T>(dir, iter, size, length, sub,
                                                         flags.data(), rho.data(), u.data(),
                                                         variant).c_str());
    } else {
        std::printf("%s\n", lbm::write_netcdf<T>(dir, rank, iter, size, length, flags.data(),
                                                  rho.data(), u.data(), variant).c_str());
    }
    return 0;
}

int main(int argc, char** argv)
{
    if (argc != 13 && argc != 16)
        return 1;
    const bool global = argc == 16;
    return std::string(argv[1]) == "f64" ? run<double>(argv, global) : run<float>(argv, global);
}
