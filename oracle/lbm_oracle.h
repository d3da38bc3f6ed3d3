/*
 * CPU oracle for the D3Q19 alpha/beta A-A step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker for the HIP product path. Only the test suite,
 * the smoke check and the benchmark's CPU-baseline leg load it. The product
 * (cavityflow_cuda_amd) never links, imports or calls it.
 *
 * It restates, in plain C, the reference CPU solver of saleemayman/cavityflow_cuda
 * operation for operation (same accumulation order, same constant folding, same
 * density/velocity output quirks), so that for T = float and T = double its
 * results are bit-identical to the reference CLbm{Init,Alpha,Beta}CPU kernels
 * when both are compiled without FMA contraction.
 *
 * Memory layout is the reference's: populations dd[l * N_h + id], velocities
 * u[c * N_h + id], densities rho[id], flags flags[id] (int32 Flag values),
 * id = z * nx_h * ny_h + y * nx_h + x (x fastest), N_h = nx_h * ny_h * nz_h.
 *
 * Pinning (see tests/test_oracle.py):
 *   - every function is checked bit for bit against golden vectors written by
 *     tests/golden/make_golden.py from the compiled reference kernels;
 *   - closed forms: rest equilibrium equals the lattice weights; tau = 1 makes
 *     FLUID post-collision values equal the equilibrium; all-GHOST beta is the
 *     pure neighbour shift; FLUID collision conserves mass and momentum.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORACLE_OBSTACLE = 1,
    ORACLE_FLUID = 2,
    ORACLE_VELOCITY_INJECTION = 4,
    ORACLE_GHOST_LAYER = 8
};

/* Flag field of CLbmInitCPU::setFlags (CLbmInitCPU.cpp:37-75).
 * order = 0: x fastest, the layout every kernel reads (lbm_init.cu:80-93).
 * order = 1: the reference CPU enumeration, x outermost and z innermost,
 *            stored in that order (z fastest), reproduced for pinning only. */
void oracle_set_flags(int32_t* flags, const int size_h[3], const int bc[6], int order);

/* T = float (suffix f) and T = double (suffix d). */
#define ORACLE_DECLARE(T, S)                                                          \
    void oracle_init_##S(T* dd, T* rho, T* u, const int size_h[3], int store_rho,     \
                         int store_u);                                                \
    void oracle_alpha_##S(T* dd, const int32_t* flags, T* rho, T* u, T inv_tau,       \
                          T lid_velocity, const T g[3], const int size_h[3],          \
                          const int origin[3], const int box[3], int store_rho,       \
                          int store_u, int nthreads);                                 \
    void oracle_beta_##S(T* dd, const int32_t* flags, T* rho, T* u, T inv_tau,        \
                         T lid_velocity, const int size_h[3], const int origin[3],    \
                         const int box[3], int store_rho, int store_u, int nthreads); \
    void oracle_get_box_##S(const T* field, int ncomp, const int size_h[3],           \
                            const int origin[3], const int box[3], T* dst);           \
    void oracle_set_box_##S(T* field, int ncomp, const int size_h[3],                 \
                            const int origin[3], const int box[3], const T* src,      \
                            uint32_t comp_mask);                                      \
    void oracle_moments_##S(const T* dd, const int size_h[3], const int origin[3],    \
                            const int box[3], int layout, T* dst);                    \
    int oracle_solver_params_##S(T length_x, int size_x, T length_max,                \
                                 const T velocity[3], const T acceleration[3],        \
                                 T viscosity, T timestep, T max_velocity,             \
                                 T max_acceleration, T out_vel_dimless[3],            \
                                 T out_acc_dimless[3], T* out_visc_dimless,           \
                                 T* out_tau, T* out_inv_tau, T* out_timestep);

ORACLE_DECLARE(float, f)
ORACLE_DECLARE(double, d)
#undef ORACLE_DECLARE

/* Lattice-vector mask of setDensityDistributions(origin, size, direction, src)
 * (CLbmSolverCPU.cpp:562-620): bit l set iff norm(direction) . c_l > 0. */
uint32_t oracle_direction_mask(int direction);

int oracle_build_id(void);

#ifdef __cplusplus
}
#endif

#endif
