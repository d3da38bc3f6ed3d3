// C++ host surface of the HIP solver: the CLbmSolver<T> operator interface of
// saleemayman/cavityflow_cuda (CLbmSolver.hpp:68-88, CLbmSolverGPU.cuh:56-97)
// over the C ABI of lbm_abi.h. Header-only; link with -lcavityflow_hip.
//
// Method names, argument order and meaning follow the reference:
//   - boxes are (origin, size) in halo coordinates;
//   - the argument-less overloads act on the whole halo-inclusive domain
//     (steps) or on the interior origin (1,1,1), size = domain size (copies);
//   - host buffers are dense [component][z][y][x].
// Where the reference asserts or exits, this class throws lbm::Error (the
// C ABI status code and message).
#ifndef CAVITYFLOW_CLBMSOLVERHIP_HPP
#define CAVITYFLOW_CLBMSOLVERHIP_HPP

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "lbm_abi.h"

namespace lbm {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status)
{
    if (status != LBM_OK)
        throw Error(status, std::string(lbm_last_error()) + " (" + lbm_status_string(status) + ")");
}

// Stand-in for CVector<3, int>: three ints with operator[] and elements().
struct Vec3i {
    int data[3];
    Vec3i(int x = 0, int y = 0, int z = 0) : data{x, y, z} {}
    int& operator[](int i) { return data[i]; }
    const int& operator[](int i) const { return data[i]; }
    long elements() const { return (long)data[0] * data[1] * data[2]; }
};

// enum Flag / Direction (common.h:54-71)
enum Flag { OBSTACLE = LBM_OBSTACLE, FLUID = LBM_FLUID,
            VELOCITY_INJECTION = LBM_VELOCITY_INJECTION, GHOST_LAYER = LBM_GHOST_LAYER };
enum Direction { LEFT = LBM_LEFT, RIGHT = LBM_RIGHT, BOTTOM = LBM_BOTTOM, TOP = LBM_TOP,
                 BACK = LBM_BACK, FRONT = LBM_FRONT };

// Derived quantities of the CLbmSolver constructor (CLbmSolver.cpp:28-281).
struct SolverParameters {
    double tau, tauInv, timestep, viscosity, viscosityDimLess;
    double velocityDimLess[3], accelerationDimLess[3];
};

inline SolverParameters computeParameters(const lbm_physics& phys, int precision)
{
    lbm_derived d;
    check(lbm_compute_parameters(&phys, precision, &d));
    SolverParameters p;
    p.tau = d.tau;
    p.tauInv = d.inv_tau;
    p.timestep = d.timestep;
    p.viscosity = d.viscosity;
    p.viscosityDimLess = d.viscosity_dimless;
    for (int i = 0; i < 3; i++) {
        p.velocityDimLess[i] = d.velocity_dimless[i];
        p.accelerationDimLess[i] = d.acceleration_dimless[i];
    }
    return p;
}

// T is the host element type of population / velocity / density buffers:
// double for LBM_F64 storage, float for LBM_F32, LBM_F16S, LBM_BF16S and
// LBM_F16C.
template <typename T>
class CLbmSolverHIP {
    static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                  "T must be float or double");

public:
    // size: interior cells (CDomain::getSize()); boundaryConditions: six Flags
    // (LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT).
    CLbmSolverHIP(int device, Vec3i size, const std::vector<Flag>& boundaryConditions,
                  const SolverParameters& derived, bool storeDensities, bool storeVelocities,
                  int precision = std::is_same<T, double>::value ? LBM_F64 : LBM_F32)
        : size_(size)
    {
        if (boundaryConditions.size() != 6)
            throw Error(LBM_ERR_ARG, "six boundary conditions are required");
        if ((precision == LBM_F64) != std::is_same<T, double>::value)
            throw Error(LBM_ERR_ARG, "T must be double exactly for LBM_F64 storage");
        lbm_params p = {};
        for (int i = 0; i < 3; i++) {
            p.size[i] = size[i];
            p.gravitation[i] = derived.accelerationDimLess[i];
        }
        for (int i = 0; i < 6; i++)
            p.bc[i] = boundaryConditions[i];
        p.precision = precision;
        p.inv_tau = derived.tauInv;
        p.lid_velocity = derived.velocityDimLess[0];
        p.store_densities = storeDensities;
        p.store_velocities = storeVelocities;
        p.device = device;
        check(lbm_create(&p, &s_));
    }
    ~CLbmSolverHIP() { lbm_destroy(s_); }
    CLbmSolverHIP(const CLbmSolverHIP&) = delete;
    CLbmSolverHIP& operator=(const CLbmSolverHIP&) = delete;

    lbm_solver* handle() { return s_; }
    Vec3i getSize() const { return size_; }
    Vec3i getSizeWithHalo() const { return Vec3i(size_[0] + 2, size_[1] + 2, size_[2] + 2); }

    // simulationStepAlpha/Beta (CLbmSolver.hpp:68-71, CLbmSolverGPU.cuh:65-72)
    void simulationStepAlpha(hipStream_t stream = nullptr)
    {
        check(lbm_simulation_step_alpha(s_, nullptr, nullptr, stream));
    }
    void simulationStepAlpha(Vec3i origin, Vec3i size, hipStream_t stream = nullptr)
    {
        check(lbm_simulation_step_alpha(s_, origin.data, size.data, stream));
    }
    void simulationStepBeta(hipStream_t stream = nullptr)
    {
        check(lbm_simulation_step_beta(s_, nullptr, nullptr, stream));
    }
    void simulationStepBeta(Vec3i origin, Vec3i size, hipStream_t stream = nullptr)
    {
        check(lbm_simulation_step_beta(s_, origin.data, size.data, stream));
    }

    // get/setDensityDistributions (CLbmSolver.hpp:73-76, CLbmSolverCPU.hpp:80)
    // and their stream overloads (CLbmSolverGPU.cuh:73-82; nullptr = the
    // default stream). With host buffers every call returns with its copy
    // complete, ordered after the work already queued on the stream.
    void getDensityDistributions(Vec3i& origin, Vec3i& size, T* dst, hipStream_t stream = nullptr)
    {
        check(lbm_get_density_distributions(s_, origin.data, size.data, dst, 0, stream));
    }
    void getDensityDistributions(T* dst, hipStream_t stream = nullptr)
    {
        check(lbm_get_density_distributions(s_, nullptr, nullptr, dst, 0, stream));
    }
    void setDensityDistributions(Vec3i& origin, Vec3i& size, T* src, hipStream_t stream = nullptr)
    {
        check(lbm_set_density_distributions(s_, origin.data, size.data, src, 0, stream));
    }
    void setDensityDistributions(Vec3i& origin, Vec3i& size, Direction direction, T* src,
                                 hipStream_t stream = nullptr)
    {
        check(lbm_set_density_distributions_dir(s_, origin.data, size.data, direction, src, 0,
                                                stream));
    }
    void setDensityDistributions(T* src, hipStream_t stream = nullptr)
    {
        check(lbm_set_density_distributions(s_, nullptr, nullptr, src, 0, stream));
    }

    // Device-buffer forms of the population copies (asynchronous on the
    // stream, as the reference GPU solver's device-side copies are): the
    // halo exchange of a multi-GPU run passes its send / receive buffers here.
    void getDensityDistributionsDevice(Vec3i& origin, Vec3i& size, T* dev_dst,
                                       hipStream_t stream)
    {
        check(lbm_get_density_distributions(s_, origin.data, size.data, dev_dst, 1, stream));
    }
    void setDensityDistributionsDevice(Vec3i& origin, Vec3i& size, Direction direction,
                                       const T* dev_src, hipStream_t stream)
    {
        check(lbm_set_density_distributions_dir(s_, origin.data, size.data, direction, dev_src,
                                                1, stream));
    }

    // The x-face exchange inside the step kernels (no reference counterpart):
    // device buffers of lbm_face_elements(LBM_LEFT) values of T each, nullptr
    // for a side without a neighbour. setStepPack: the steps pack the 5
    // populations per face and phase that lbm_pack_face packs; setStepUnpack:
    // the steps take the received ones, with the results of lbm_unpack_face
    // before each step (exact where x is the only split axis).
    void setStepPack(T* left, T* right) { check(lbm_set_step_pack(s_, left, right)); }
    void setStepUnpack(const T* left, const T* right)
    {
        check(lbm_set_step_unpack(s_, left, right));
    }

    // get/setFlags (CLbmSolver.hpp:77-80)
    void getFlags(Vec3i& origin, Vec3i& size, Flag* dst)
    {
        check(lbm_get_flags(s_, origin.data, size.data, reinterpret_cast<int32_t*>(dst)));
    }
    void getFlags(Flag* dst) { check(lbm_get_flags(s_, nullptr, nullptr, reinterpret_cast<int32_t*>(dst))); }
    void setFlags(Vec3i& origin, Vec3i& size, Flag* src)
    {
        check(lbm_set_flags(s_, origin.data, size.data, reinterpret_cast<const int32_t*>(src)));
    }
    void setFlags(Flag* src) { check(lbm_set_flags(s_, nullptr, nullptr, reinterpret_cast<const int32_t*>(src))); }

    // get/setVelocities, get/setDensities (CLbmSolver.hpp:81-88)
    void getVelocities(Vec3i& origin, Vec3i& size, T* dst)
    {
        check(lbm_get_velocities(s_, origin.data, size.data, dst, 0, nullptr));
    }
    void getVelocities(T* dst) { check(lbm_get_velocities(s_, nullptr, nullptr, dst, 0, nullptr)); }
    void setVelocities(Vec3i& origin, Vec3i& size, T* src)
    {
        check(lbm_set_velocities(s_, origin.data, size.data, src, 0, nullptr));
    }
    void setVelocities(T* src) { check(lbm_set_velocities(s_, nullptr, nullptr, src, 0, nullptr)); }
    void getDensities(Vec3i& origin, Vec3i& size, T* dst)
    {
        check(lbm_get_densities(s_, origin.data, size.data, dst, 0, nullptr));
    }
    void getDensities(T* dst) { check(lbm_get_densities(s_, nullptr, nullptr, dst, 0, nullptr)); }
    void setDensities(Vec3i& origin, Vec3i& size, T* src)
    {
        check(lbm_set_densities(s_, origin.data, size.data, src, 0, nullptr));
    }
    void setDensities(T* src) { check(lbm_set_densities(s_, nullptr, nullptr, src, 0, nullptr)); }

    // CController::setDrivenCavitySzenario (CController.cpp:1313-1345)
    void setDrivenCavitySzenario()
    {
        Vec3i origin(1, size_[1], 1);   // top interior layer y = size_h - 2
        Vec3i box(size_[0], 1, size_[2]);
        std::vector<Flag> src((size_t)box.elements(), VELOCITY_INJECTION);
        setFlags(origin, box, src.data());
    }

private:
    lbm_solver* s_ = nullptr;
    Vec3i size_;
};

}  // namespace lbm

#endif
