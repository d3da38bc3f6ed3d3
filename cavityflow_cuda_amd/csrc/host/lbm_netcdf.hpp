// netCDF output, the reference's active writer (CLbmVisualizationNetCDF.cpp:
// 92-228), written without the netCDF library in the classic format:
// big-endian header, then each variable's values. One file per rank (the
// serial build) or one for the whole domain (the parallel build).
//
// Layout (as the reference defines it, CLbmVisualizationNetCDF.cpp:92-155):
//   dimensions x, y, z (defined in that order: ids 0, 1, 2);
//   variables x(x), y(y), z(z) of type T with cell-centre coordinates
//   (i + 0.5) * length / size computed in T; flags(z, y, x) int;
//   densities, velocities_x, velocities_y, velocities_z (z, y, x) of type T;
//   each with a "units" text attribute; no global attributes, no records.
// The data section starts right after the header (4-byte alignment), the
// variables follow one another in definition order, each padded to 4 bytes.
// Files whose variable offsets need more than 31 bits use the 64-bit-offset
// variant of the format (version byte 2); a variable of 4 GiB or more cannot
// be stored in either and is rejected.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <typeinfo>
#include <vector>

namespace lbm {
namespace netcdf {

enum : int32_t { kChar = 2, kInt = 4, kFloat = 5, kDouble = 6 };
enum : int32_t { kDimension = 10, kVariable = 11, kAttribute = 12 };

class Header {
public:
    explicit Header(bool offset64) : offset64_(offset64) {}
    void i32(int32_t v)
    {
        const uint32_t u = (uint32_t)v;
        const uint8_t b[4] = {(uint8_t)(u >> 24), (uint8_t)(u >> 16), (uint8_t)(u >> 8),
                              (uint8_t)u};
        bytes_.insert(bytes_.end(), b, b + 4);
    }
    void offset(uint64_t v)
    {
        if (offset64_)
            i32((int32_t)(uint32_t)(v >> 32));
        i32((int32_t)(uint32_t)v);
    }
    void text(const std::string& s)          // nelems, characters, zero padding
    {
        i32((int32_t)s.size());
        bytes_.insert(bytes_.end(), s.begin(), s.end());
        while (bytes_.size() % 4)
            bytes_.push_back(0);
    }
    void raw(const char* s, size_t n) { bytes_.insert(bytes_.end(), s, s + n); }
    const std::vector<uint8_t>& bytes() const { return bytes_; }
    size_t size() const { return bytes_.size(); }

private:
    bool offset64_;
    std::vector<uint8_t> bytes_;
};

struct Var {
    std::string name;
    std::vector<int32_t> dims;   // dimension ids, slowest first
    int32_t type;
    std::string units;
    uint64_t bytes;              // unpadded size of the values
};

inline uint64_t padded(uint64_t n) { return (n + 3) & ~(uint64_t)3; }

inline std::vector<uint8_t> header(const int size[3], const std::vector<Var>& vars, bool offset64)
{
    const char* dim_names[3] = {"x", "y", "z"};
    Header h(offset64);
    h.raw("CDF", 3);
    const char version = offset64 ? 2 : 1;
    h.raw(&version, 1);
    h.i32(0);                                            // numrecs
    h.i32(kDimension);
    h.i32(3);
    for (int d = 0; d < 3; d++) {
        h.text(dim_names[d]);
        h.i32(size[d]);
    }
    h.i32(0);                                            // no global attributes
    h.i32(0);
    h.i32(kVariable);
    h.i32((int32_t)vars.size());
    // pass 1 sizes the header (offsets have a fixed width), pass 2 fills them
    uint64_t begin = 0;
    for (int pass = 0; pass < 2; pass++) {
        Header v(offset64);
        uint64_t at = begin;
        for (const Var& var : vars) {
            v.text(var.name);
            v.i32((int32_t)var.dims.size());
            for (int32_t d : var.dims)
                v.i32(d);
            v.i32(kAttribute);
            v.i32(1);
            v.text("units");
            v.i32(kChar);
            v.text(var.units);
            v.i32(var.type);
            v.i32((int32_t)(uint32_t)padded(var.bytes));
            v.offset(at);
            at += padded(var.bytes);
        }
        if (pass == 0) {
            begin = h.size() + v.size();
        } else {
            h.raw((const char*)v.bytes().data(), v.size());
        }
    }
    return h.bytes();
}

template <class V>
void put_be(std::ofstream& f, const V* values, uint64_t count)
{
    std::vector<uint8_t> buf;
    const uint64_t chunk = 1u << 20;
    for (uint64_t first = 0; first < count; first += chunk) {
        const uint64_t n = std::min(chunk, count - first);
        buf.resize(n * sizeof(V));
        for (uint64_t i = 0; i < n; i++) {
            uint8_t b[sizeof(V)];
            std::memcpy(b, &values[first + i], sizeof(V));
            for (size_t k = 0; k < sizeof(V); k++)
                buf[i * sizeof(V) + k] = b[sizeof(V) - 1 - k];
        }
        f.write((const char*)buf.data(), (std::streamsize)buf.size());
    }
    static const char zeros[4] = {0, 0, 0, 0};
    f.write(zeros, (std::streamsize)(padded(count * sizeof(V)) - count * sizeof(V)));
}

}  // namespace netcdf

// One netCDF file of interior fields flags / rho [z][y][x] and u [3][z][y][x]
// over `size` cells; the coordinate variables hold the cell centres
// ((T)i + 0.5) * cell_length / (T)cell_size (CLbmVisualizationNetCDF.cpp:
// 166-203: the subdomain's length and size, also for the whole-domain file of
// the parallel build). variant: 0 chooses the classic or 64-bit-offset form
// by size, 1 / 2 force it.
template <class T>
void write_netcdf_file(const std::string& path, const int size[3], const T cell_length[3],
                       const int cell_size[3], const int32_t* flags, const T* rho, const T* u,
                       int variant = 0)
{
    using namespace netcdf;
    const int32_t tt = typeid(T) == typeid(double) ? kDouble : kFloat;
    const uint64_t cells = (uint64_t)size[0] * size[1] * size[2];
    const std::vector<int32_t> zyx = {2, 1, 0};
    std::vector<Var> vars = {
        {"x", {0}, tt, "meters", (uint64_t)size[0] * sizeof(T)},
        {"y", {1}, tt, "meters", (uint64_t)size[1] * sizeof(T)},
        {"z", {2}, tt, "meters", (uint64_t)size[2] * sizeof(T)},
        {"flags", zyx, kInt, "flag", cells * 4},
        {"densities", zyx, tt, "kg/m^3", cells * sizeof(T)},
        {"velocities_x", zyx, tt, "m/s", cells * sizeof(T)},
        {"velocities_y", zyx, tt, "m/s", cells * sizeof(T)},
        {"velocities_z", zyx, tt, "m/s", cells * sizeof(T)},
    };
    // classic: offsets and variable sizes below 2^31 - 4; 64-bit-offset
    // variant: variable sizes below 2^32 - 4 (the header is below 1 KiB)
    uint64_t data = 0, last_begin = 0, largest = 0;
    for (const Var& v : vars) {
        if (padded(v.bytes) > (1ull << 32) - 4)
            throw std::runtime_error("netCDF: a variable of 4 GiB or more does not fit the "
                                     "classic format");
        last_begin = data;
        largest = std::max(largest, padded(v.bytes));
        data += padded(v.bytes);
    }
    const bool offset64 = variant ? variant == 2
                                  : last_begin + 1024 > 0x7fffffffull || largest > (1ull << 31) - 4;
    const std::vector<uint8_t> head = header(size, vars, offset64);

    std::ofstream f(path.c_str(), std::ios::out | std::ios::binary | std::ios::trunc);
    if (!f)
        throw std::runtime_error("cannot open " + path);
    f.write((const char*)head.data(), (std::streamsize)head.size());
    for (int d = 0; d < 3; d++) {
        std::vector<T> c(size[d]);
        for (int i = 0; i < size[d]; i++)
            c[i] = ((T)i + (T)0.5) * cell_length[d] / (T)cell_size[d];
        put_be(f, c.data(), (uint64_t)size[d]);
    }
    put_be(f, flags, cells);
    put_be(f, rho, cells);
    for (int k = 0; k < 3; k++)
        put_be(f, u + k * cells, cells);
    if (!f)
        throw std::runtime_error("write failed: " + path);
}

// visualization_<rank>_<iteration>.nc in `dir` (the serial build's file name,
// CLbmVisualizationNetCDF.cpp:55-56) for one subdomain of `size` cells and
// physical `length`.
template <class T>
std::string write_netcdf(const std::string& dir, int rank, int iteration, const int size[3],
                         const T length[3], const int32_t* flags, const T* rho, const T* u,
                         int variant = 0)
{
    std::stringstream name;
    name << dir << "/visualization_" << rank << "_" << iteration << ".nc";
    write_netcdf_file(name.str(), size, length, size, flags, rho, u, variant);
    return name.str();
}

// visualization_<iteration>.nc in `dir` (the parallel build's file,
// CLbmVisualizationNetCDF.cpp:52, 107-111): the whole domain of `size`
// cells, each subdomain's fields at its global origin; coordinates from the
// subdomain's length and size.
template <class T>
std::string write_netcdf_global(const std::string& dir, int iteration, const int size[3],
                                const T sub_length[3], const int sub_size[3],
                                const int32_t* flags, const T* rho, const T* u, int variant = 0)
{
    std::stringstream name;
    name << dir << "/visualization_" << iteration << ".nc";
    write_netcdf_file(name.str(), size, sub_length, sub_size, flags, rho, u, variant);
    return name.str();
}

}  // namespace lbm
