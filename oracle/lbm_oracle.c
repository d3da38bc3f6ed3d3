/*
 * CPU oracle for the D3Q19 alpha/beta A-A step -- TEST INFRASTRUCTURE ONLY.
 * Loaded only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg, as the checker. Build: oracle/Makefile (-O3 -fopenmp -ffp-contract=off).
 */
#include "lbm_oracle.h"

#include <float.h>
#include <math.h>

/* D3Q19 velocity set E0..E18, common.h:29-47 (lbm_units[], main.cpp:34-40). */
static const int oracle_c[19][3] = {
    {1, 0, 0},  {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0}, {1, 0, 1},  {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {0, 1, 1},   {0, -1, -1},
    {0, 1, -1}, {0, -1, 1}, {0, 0, 1},  {0, 0, -1}, {0, 0, 0}};

/* constants.h:24-28; the two defaults are double literals in the reference */
#define ORACLE_REYNOLDS_DEFAULT (1000.0)
#define ORACLE_TAU_DEFAULT (0.85)
#define ORACLE_TAU_LOWER_LIMIT (0.51)
#define ORACLE_TAU_UPPER_LIMIT (1.9)

/* CMath<float>::sqrt = sqrtf, CMath<double>::sqrt = ::sqrt (CMath.cpp:263-271) */
static inline float oracle_sqrt_f(float a) { return sqrtf(a); }
static inline double oracle_sqrt_d(double a) { return sqrt(a); }

#define T float
#define S f
#define ORACLE_EPS FLT_EPSILON
#include "lbm_oracle_impl.h"
#undef ORACLE_EPS
#undef S
#undef T

#define T double
#define S d
#define ORACLE_EPS DBL_EPSILON
#include "lbm_oracle_impl.h"
#undef ORACLE_EPS
#undef S
#undef T

/* CLbmInitCPU::setFlags, CLbmInitCPU.cpp:37-75: the first OBSTACLE wins,
 * GHOST_LAYER can be overwritten by a later face. */
void oracle_set_flags(int32_t* flags, const int size_h[3], const int bc[6], int order)
{
    int x, y, z;
    long id = 0;
    for (x = 0; x < size_h[0]; x++)
        for (y = 0; y < size_h[1]; y++)
            for (z = 0; z < size_h[2]; z++) {
                int32_t flag = ORACLE_FLUID;
                if (x == 0)
                    flag = bc[0];
                if (x == size_h[0] - 1 && flag != ORACLE_OBSTACLE)
                    flag = bc[1];
                if (y == 0 && flag != ORACLE_OBSTACLE)
                    flag = bc[2];
                if (y == size_h[1] - 1 && flag != ORACLE_OBSTACLE)
                    flag = bc[3];
                if (z == 0 && flag != ORACLE_OBSTACLE)
                    flag = bc[4];
                if (z == size_h[2] - 1 && flag != ORACLE_OBSTACLE)
                    flag = bc[5];
                if (order == 1)
                    flags[id++] = flag;
                else
                    flags[((long)z * size_h[1] + y) * size_h[0] + x] = flag;
            }
}

/* CLbmSolverCPU::setDensityDistributions(origin, size, direction, src):
 * norm per direction (CLbmSolverCPU.cpp:564-586), lattice vector l copied iff
 * norm . c_l > 0 (:616-619). */
uint32_t oracle_direction_mask(int direction)
{
    int norm[3] = {0, 0, 0};
    uint32_t mask = 0;
    int l;
    switch (direction) {
    case 0: norm[0] = 1; break;  /* LEFT */
    case 1: norm[0] = -1; break; /* RIGHT */
    case 2: norm[1] = 1; break;  /* BOTTOM */
    case 3: norm[1] = -1; break; /* TOP */
    case 4: norm[2] = 1; break;  /* BACK */
    case 5: norm[2] = -1; break; /* FRONT */
    default: return 0;
    }
    for (l = 0; l < 19; l++)
        if (norm[0] * oracle_c[l][0] + norm[1] * oracle_c[l][1] + norm[2] * oracle_c[l][2] > 0)
            mask |= 1u << l;
    return mask;
}

int oracle_build_id(void) { return 1; }
