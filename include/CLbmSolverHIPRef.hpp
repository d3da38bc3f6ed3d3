// The HIP solver as a CLbmSolver<T> of saleemayman/cavityflow_cuda: the
// subclass a maintainer adds to the reference tree (as src/CLbmSolverHIP.hpp)
// in place of CLbmSolverGPU<T>. It derives from the reference's own abstract
// solver (CLbmSolver.hpp:30-89), so CLbmSolver's constructor performs the
// tau / time-step adaptation exactly as for the CUDA solver, and forwards every
// method to the C ABI of lbm_abi.h (link -lcavityflow_hip).
//
// Surface: the virtual methods of CLbmSolver.hpp:68-88 plus the stream
// overloads of CLbmSolverGPU.cuh:65-94, with hipStream_t* in place of
// cudaStream_t* (NULL = the default stream). Host buffers are dense
// [component][z][y][x], as in the reference. Like the reference's
// GPU_ERROR_CHECK (common.h:73-90) a failing call prints the error and exits.
//
// Compiled against the reference's headers by tests/test_reference_binding.py
// (oracle/Makefile target `binding`).
#ifndef CAVITYFLOW_CLBMSOLVERHIP_REF_HPP
#define CAVITYFLOW_CLBMSOLVERHIP_REF_HPP

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "CLbmSolver.hpp"
#include "lbm_abi.h"

template <typename T>
class CLbmSolverHIP : public CLbmSolver<T>
{
private:
    lbm_solver* solver;

    static void check(int status, const char* call)
    {
        if (status != LBM_OK) {
            std::fprintf(stderr, "%s failed: %s (%s)\n", call, lbm_last_error(),
                         lbm_status_string(status));
            std::exit(EXIT_FAILURE);
        }
    }
    static hipStream_t stream_of(hipStream_t* stream) { return stream ? *stream : NULL; }

    // CVector<3, int> -> int[3]
    struct Box {
        int origin[3], size[3];
        Box(CVector<3, int>& o, CVector<3, int>& s)
        {
            for (int i = 0; i < 3; i++) {
                origin[i] = o[i];
                size[i] = s[i];
            }
        }
    };

public:
    CLbmSolverHIP(int id, CDomain<T>& domain, std::vector<Flag> boundaryConditions,
                  CConfiguration<T>* configuration)
        : CLbmSolver<T>(id, domain, boundaryConditions, configuration), solver(NULL)
    {
        int devices = 0;
        check(lbm_device_count(&devices), "lbm_device_count");
        if (devices < 1) {
            std::fprintf(stderr, "CLbmSolverHIP: no GPU\n");
            std::exit(EXIT_FAILURE);
        }
        CVector<3, int> size = domain.getSize();
        lbm_params p = {};
        for (int i = 0; i < 3; i++) {
            p.size[i] = size[i];
            p.gravitation[i] = this->accelerationDimLess[i];
        }
        for (int i = 0; i < 6; i++)
            p.bc[i] = this->boundaryConditions[i];   // the Flag values are the ABI's
        p.precision = sizeof(T) == sizeof(double) ? LBM_F64 : LBM_F32;
        p.inv_tau = this->tauInv;
        p.lid_velocity = this->velocityDimLess[0];
        p.store_densities = this->storeDensities ? 1 : 0;
        p.store_velocities = this->storeVelocities ? 1 : 0;
        p.device = this->id % devices;               // CLbmSolverGPU.cu:43-45
        check(lbm_create(&p, &solver), "lbm_create");
    }
    ~CLbmSolverHIP() { lbm_destroy(solver); }

    lbm_solver* handle() { return solver; }

    // simulationStepAlpha / simulationStepBeta (CLbmSolverGPU.cuh:65-72)
    void simulationStepAlpha(hipStream_t* stream)
    {
        check(lbm_simulation_step_alpha(solver, NULL, NULL, stream_of(stream)), "step alpha");
    }
    void simulationStepAlpha() { simulationStepAlpha((hipStream_t*)NULL); }
    void simulationStepAlpha(CVector<3, int> origin, CVector<3, int> size, hipStream_t* stream)
    {
        Box b(origin, size);
        check(lbm_simulation_step_alpha(solver, b.origin, b.size, stream_of(stream)),
              "step alpha");
    }
    void simulationStepAlpha(CVector<3, int> origin, CVector<3, int> size)
    {
        simulationStepAlpha(origin, size, NULL);
    }
    void simulationStepBeta(hipStream_t* stream)
    {
        check(lbm_simulation_step_beta(solver, NULL, NULL, stream_of(stream)), "step beta");
    }
    void simulationStepBeta() { simulationStepBeta((hipStream_t*)NULL); }
    void simulationStepBeta(CVector<3, int> origin, CVector<3, int> size, hipStream_t* stream)
    {
        Box b(origin, size);
        check(lbm_simulation_step_beta(solver, b.origin, b.size, stream_of(stream)),
              "step beta");
    }
    void simulationStepBeta(CVector<3, int> origin, CVector<3, int> size)
    {
        simulationStepBeta(origin, size, NULL);
    }

    // get/setDensityDistributions (CLbmSolverGPU.cuh:73-82). With host
    // buffers each call returns with its copy complete.
    void getDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size, T* dst,
                                 hipStream_t* stream)
    {
        Box b(origin, size);
        check(lbm_get_density_distributions(solver, b.origin, b.size, dst, 0, stream_of(stream)),
              "getDensityDistributions");
    }
    void getDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size, T* dst)
    {
        getDensityDistributions(origin, size, dst, NULL);
    }
    void getDensityDistributions(T* dst, hipStream_t* stream)
    {
        check(lbm_get_density_distributions(solver, NULL, NULL, dst, 0, stream_of(stream)),
              "getDensityDistributions");
    }
    void getDensityDistributions(T* dst) { getDensityDistributions(dst, NULL); }
    void setDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size,
                                 Direction direction, T* src, hipStream_t* stream)
    {
        Box b(origin, size);
        check(lbm_set_density_distributions_dir(solver, b.origin, b.size, direction, src, 0,
                                                stream_of(stream)),
              "setDensityDistributions");
    }
    void setDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size,
                                 Direction direction, T* src)
    {
        setDensityDistributions(origin, size, direction, src, NULL);
    }
    void setDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size, T* src,
                                 hipStream_t* stream)
    {
        Box b(origin, size);
        check(lbm_set_density_distributions(solver, b.origin, b.size, src, 0, stream_of(stream)),
              "setDensityDistributions");
    }
    void setDensityDistributions(CVector<3, int>& origin, CVector<3, int>& size, T* src)
    {
        setDensityDistributions(origin, size, src, NULL);
    }
    void setDensityDistributions(T* src, hipStream_t* stream)
    {
        check(lbm_set_density_distributions(solver, NULL, NULL, src, 0, stream_of(stream)),
              "setDensityDistributions");
    }
    void setDensityDistributions(T* src) { setDensityDistributions(src, NULL); }

    // get/setFlags, get/setVelocities, get/setDensities (CLbmSolver.hpp:77-88)
    void getFlags(CVector<3, int>& origin, CVector<3, int>& size, Flag* dst)
    {
        Box b(origin, size);
        check(lbm_get_flags(solver, b.origin, b.size, reinterpret_cast<int32_t*>(dst)),
              "getFlags");
    }
    void getFlags(Flag* dst)
    {
        check(lbm_get_flags(solver, NULL, NULL, reinterpret_cast<int32_t*>(dst)), "getFlags");
    }
    void setFlags(CVector<3, int>& origin, CVector<3, int>& size, Flag* src)
    {
        Box b(origin, size);
        check(lbm_set_flags(solver, b.origin, b.size, reinterpret_cast<const int32_t*>(src)),
              "setFlags");
    }
    void setFlags(Flag* src)
    {
        check(lbm_set_flags(solver, NULL, NULL, reinterpret_cast<const int32_t*>(src)),
              "setFlags");
    }
    void getVelocities(CVector<3, int>& origin, CVector<3, int>& size, T* dst)
    {
        Box b(origin, size);
        check(lbm_get_velocities(solver, b.origin, b.size, dst, 0, NULL), "getVelocities");
    }
    void getVelocities(T* dst)
    {
        check(lbm_get_velocities(solver, NULL, NULL, dst, 0, NULL), "getVelocities");
    }
    void setVelocities(CVector<3, int>& origin, CVector<3, int>& size, T* src)
    {
        Box b(origin, size);
        check(lbm_set_velocities(solver, b.origin, b.size, src, 0, NULL), "setVelocities");
    }
    void setVelocities(T* src)
    {
        check(lbm_set_velocities(solver, NULL, NULL, src, 0, NULL), "setVelocities");
    }
    void getDensities(CVector<3, int>& origin, CVector<3, int>& size, T* dst)
    {
        Box b(origin, size);
        check(lbm_get_densities(solver, b.origin, b.size, dst, 0, NULL), "getDensities");
    }
    void getDensities(T* dst)
    {
        check(lbm_get_densities(solver, NULL, NULL, dst, 0, NULL), "getDensities");
    }
    void setDensities(CVector<3, int>& origin, CVector<3, int>& size, T* src)
    {
        Box b(origin, size);
        check(lbm_set_densities(solver, b.origin, b.size, src, 0, NULL), "setDensities");
    }
    void setDensities(T* src)
    {
        check(lbm_set_densities(solver, NULL, NULL, src, 0, NULL), "setDensities");
    }
};

#endif
