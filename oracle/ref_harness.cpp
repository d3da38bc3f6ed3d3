// Harness around the reference's own CPU kernels -- TEST INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile together with the unmodified reference sources
// under /root/reference/src (CLbm{Init,Alpha,Beta}CPU.cpp, CLbmSolver.cpp,
// CConfiguration.cpp, CDomain.cpp, libmath, vendored tinyxml2) into
// oracle/_ref/libref_lbm.so. It exists to (1) write the golden vectors under
// tests/golden/ (kernels, CLbmSolver parameters, CConfiguration fields, VTK
// files of CLbmVisualizationVTK) and (2) time the reference CPU path as the
// CPU baseline. Built with USE_MPI defined, as the reference's MPI build
// (common.make:66-67), which only changes how CConfiguration reads
// subdomain-num.
// The product never loads it.
#include <cstdint>
#include <cstdio>
#include <vector>

#include "CConfiguration.hpp"
#include "CDomain.hpp"
#include "CLbmSolver.hpp"
#include "cpukernels/CLbmAlphaCPU.hpp"
#include "cpukernels/CLbmBetaCPU.hpp"
#include "cpukernels/CLbmInitCPU.hpp"
#include "libvis/CLbmVisualizationVTK.hpp"

static_assert(sizeof(Flag) == sizeof(int32_t), "reference Flag must be 32-bit");

// The reference declares lbm_units extern in common.h and defines it in
// main.cpp, which is not part of this build.
CVector<3, int> lbm_units[] = {E0,  E1,  E2,  E3,  E4,  E5,  E6,  E7,  E8, E9,
                               E10, E11, E12, E13, E14, E15, E16, E17, E18};

namespace {

std::vector<Flag> to_bc(const int bc[6])
{
    std::vector<Flag> v(6);
    for (int i = 0; i < 6; i++)
        v[i] = static_cast<Flag>(bc[i]);
    return v;
}

CVector<3, int> v3(const int a[3]) { return CVector<3, int>(a[0], a[1], a[2]); }

template <class T>
void init(T* dd, int32_t* flags, T* rho, T* u, const int size_h[3], const int bc[6],
          int store_rho, int store_u)
{
    CLbmInitCPU<T> k(v3(size_h), to_bc(bc));
    k.initLbm(dd, reinterpret_cast<Flag*>(flags), rho, u, store_rho != 0, store_u != 0);
}

template <class T>
void alpha(T* dd, int32_t* flags, T* rho, T* u, T inv_tau, T lid, const T g[3],
           const int size_h[3], const int origin[3], const int box[3], const int block[3],
           int store_rho, int store_u)
{
    CLbmAlphaCPU<T> k(v3(size_h), CVector<3, T>(g[0], g[1], g[2]));
    k.alphaKernelCPU(dd, reinterpret_cast<Flag*>(flags), rho, u, inv_tau, lid, v3(origin),
                     v3(box), v3(block), store_rho != 0, store_u != 0);
}

template <class T>
void beta(T* dd, int32_t* flags, T* rho, T* u, T inv_tau, T lid, const int size_h[3],
          const int origin[3], const int box[3], const int block[3], int store_rho,
          int store_u)
{
    CLbmBetaCPU<T> k(v3(size_h), CVector<3, T>(0, 0, 0));
    // CLbmSolverCPU::simulationStepBeta passes isPowerOfTwo(cells with halo)
    const int n = size_h[0] * size_h[1] * size_h[2];
    k.betaKernelCPU(dd, reinterpret_cast<Flag*>(flags), rho, u, inv_tau, lid, v3(origin),
                    v3(box), v3(block), isPowerOfTwo(n) != 0, store_rho != 0, store_u != 0);
}

// Exposes the derived quantities of the CLbmSolver constructor.
template <class T>
struct SolverProbe : public CLbmSolver<T> {
    SolverProbe(CDomain<T>& d, std::vector<Flag> bc, CConfiguration<T>* c)
        : CLbmSolver<T>(0, d, bc, c)
    {
    }
    void read(T out[12])
    {
        out[0] = this->velocityDimLess[0];
        out[1] = this->velocityDimLess[1];
        out[2] = this->velocityDimLess[2];
        out[3] = this->accelerationDimLess[0];
        out[4] = this->accelerationDimLess[1];
        out[5] = this->accelerationDimLess[2];
        out[6] = this->viscosityDimLess;
        out[7] = this->tau;
        out[8] = this->tauInv;
        out[9] = this->timestepSize;
        out[10] = this->viscosity;
        out[11] = (T)0;
    }
};

// Hands given interior flags / densities / velocities to the reference's VTK
// writer (CLbmVisualizationVTK), which reads them through these getters.
template <class T>
struct VisProbe : public CLbmSolver<T> {
    const int32_t* f;
    const T* rho;
    const T* u;
    VisProbe(CDomain<T>& d, std::vector<Flag> bc, CConfiguration<T>* c, const int32_t* f,
             const T* rho, const T* u)
        : CLbmSolver<T>(0, d, bc, c), f(f), rho(rho), u(u)
    {
    }
    int cells() { return this->getDomain()->getNumOfCells(); }
    void getFlags(Flag* dst)
    {
        for (int i = 0; i < cells(); i++)
            dst[i] = static_cast<Flag>(f[i]);
    }
    void getDensities(T* dst)
    {
        for (int i = 0; i < cells(); i++)
            dst[i] = rho[i];
    }
    void getVelocities(T* dst)
    {
        for (int i = 0; i < 3 * cells(); i++)
            dst[i] = u[i];
    }
};

template <class T>
void vtk(const char* xml_path, const char* dir, int id, int iteration, const int size[3],
         const int origin[3], const T length[3], const int32_t* flags, const T* rho, const T* u)
{
    CConfiguration<T> cfg(xml_path);
    CDomain<T> dom(id, v3(size), v3(origin), CVector<3, T>(length[0], length[1], length[2]));
    std::vector<Flag> bc(6, OBSTACLE);
    VisProbe<T> p(dom, bc, &cfg, flags, rho, u);
    CLbmVisualizationVTK<T> vis(id, 1, &p, dir);
    vis.render(iteration);
}

// Every field CConfiguration reads from the XML file, as double, in this
// order: velocity[3], acceleration[3], viscosity, max velocity, max
// acceleration, domain size[3], domain length[3], subdomains[3], CPU ratio,
// loops, timestep, do benchmark / logging / validation / visualization,
// visualization rate, alpha / beta CPU blocks[3 + 3], init / alpha / beta GPU
// grids[3 + 3 + 3]; then the visualization output directory.
template <class T>
void config(const char* xml_path, double out[44], char* vis_dir, int cap)
{
    CConfiguration<T> c(xml_path);
    int k = 0;
    for (int i = 0; i < 3; i++) out[k++] = c.velocity[i];
    for (int i = 0; i < 3; i++) out[k++] = c.acceleration[i];
    out[k++] = c.viscosity;
    out[k++] = c.maxVelocityDimLess;
    out[k++] = c.maxAccelerationDimLess;
    for (int i = 0; i < 3; i++) out[k++] = c.domainSize[i];
    for (int i = 0; i < 3; i++) out[k++] = c.domainLength[i];
    for (int i = 0; i < 3; i++) out[k++] = c.numOfSubdomains[i];
    out[k++] = c.CPUSubdomainRatio;
    out[k++] = c.loops;
    out[k++] = c.timestep;
    out[k++] = c.doBenchmark;
    out[k++] = c.doLogging;
    out[k++] = c.doValidation;
    out[k++] = c.doVisualization;
    out[k++] = c.visualizationRate;
    for (int b = 0; b < 2; b++)
        for (int i = 0; i < 3; i++) out[k++] = c.elementsPerBlock[b][i];
    for (int g = 0; g < 3; g++) {
        out[k++] = c.threadsPerBlock[g].x;
        out[k++] = c.threadsPerBlock[g].y;
        out[k++] = c.threadsPerBlock[g].z;
    }
    snprintf(vis_dir, cap, "%s", c.visualizationOutputDir.c_str());
}

template <class T>
void solver_params(const char* xml_path, T out[12])
{
    CConfiguration<T> cfg(xml_path);
    CDomain<T> dom(0, cfg.domainSize, CVector<3, int>(0), cfg.domainLength);
    std::vector<Flag> bc(6, OBSTACLE);
    SolverProbe<T> p(dom, bc, &cfg);
    p.read(out);
}

}  // namespace

extern "C" {

void ref_init_f(float* dd, int32_t* flags, float* rho, float* u, const int size_h[3],
                const int bc[6], int store_rho, int store_u)
{
    init<float>(dd, flags, rho, u, size_h, bc, store_rho, store_u);
}
void ref_init_d(double* dd, int32_t* flags, double* rho, double* u, const int size_h[3],
                const int bc[6], int store_rho, int store_u)
{
    init<double>(dd, flags, rho, u, size_h, bc, store_rho, store_u);
}
void ref_alpha_f(float* dd, int32_t* flags, float* rho, float* u, float inv_tau, float lid,
                 const float g[3], const int size_h[3], const int origin[3], const int box[3],
                 const int block[3], int store_rho, int store_u)
{
    alpha<float>(dd, flags, rho, u, inv_tau, lid, g, size_h, origin, box, block, store_rho,
                 store_u);
}
void ref_alpha_d(double* dd, int32_t* flags, double* rho, double* u, double inv_tau,
                 double lid, const double g[3], const int size_h[3], const int origin[3],
                 const int box[3], const int block[3], int store_rho, int store_u)
{
    alpha<double>(dd, flags, rho, u, inv_tau, lid, g, size_h, origin, box, block, store_rho,
                  store_u);
}
void ref_beta_f(float* dd, int32_t* flags, float* rho, float* u, float inv_tau, float lid,
                const int size_h[3], const int origin[3], const int box[3], const int block[3],
                int store_rho, int store_u)
{
    beta<float>(dd, flags, rho, u, inv_tau, lid, size_h, origin, box, block, store_rho,
                store_u);
}
void ref_beta_d(double* dd, int32_t* flags, double* rho, double* u, double inv_tau, double lid,
                const int size_h[3], const int origin[3], const int box[3], const int block[3],
                int store_rho, int store_u)
{
    beta<double>(dd, flags, rho, u, inv_tau, lid, size_h, origin, box, block, store_rho,
                 store_u);
}
void ref_solver_params_f(const char* xml_path, float out[12])
{
    solver_params<float>(xml_path, out);
}
void ref_solver_params_d(const char* xml_path, double out[12])
{
    solver_params<double>(xml_path, out);
}
void ref_vtk_f(const char* xml_path, const char* dir, int id, int iteration, const int size[3],
               const int origin[3], const float length[3], const int32_t* flags,
               const float* rho, const float* u)
{
    vtk<float>(xml_path, dir, id, iteration, size, origin, length, flags, rho, u);
}
void ref_vtk_d(const char* xml_path, const char* dir, int id, int iteration, const int size[3],
               const int origin[3], const double length[3], const int32_t* flags,
               const double* rho, const double* u)
{
    vtk<double>(xml_path, dir, id, iteration, size, origin, length, flags, rho, u);
}
void ref_config_f(const char* xml_path, double out[44], char* vis_dir, int cap)
{
    config<float>(xml_path, out, vis_dir, cap);
}
void ref_config_d(const char* xml_path, double out[44], char* vis_dir, int cap)
{
    config<double>(xml_path, out, vis_dir, cap);
}
int ref_build_id(void) { return 2; }

}  // extern "C"
