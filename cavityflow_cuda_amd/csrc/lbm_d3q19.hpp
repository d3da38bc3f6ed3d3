// D3Q19 lattice, flags and the reference's equilibrium arithmetic, shared by
// host and device code of the HIP solver.
//
// Lattice velocities E0..E18: common.h:29-47. Flags: common.h:54-61.
// Equilibrium: common.h:132-160 (incompressible form: u is the momentum j,
// never divided by rho). Every expression keeps the reference's operand order
// so that, built with -ffp-contract=off, FP32/FP64 results round exactly as
// the reference CPU kernels do.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

constexpr int kQ = 19;

enum FlagValue : uint8_t {
    kObstacle = 1,
    kFluid = 2,
    kVelocityInjection = 4,
    kGhostLayer = 8,
};

// c_l components; opposite pairs are (2k, 2k+1), 18 is rest.
__host__ __device__ constexpr int cx(int l)
{
    constexpr int v[kQ] = {1, -1, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 0, 0, 0};
    return v[l];
}
__host__ __device__ constexpr int cy(int l)
{
    constexpr int v[kQ] = {0, 0, 1, -1, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0};
    return v[l];
}
__host__ __device__ constexpr int cz(int l)
{
    constexpr int v[kQ] = {0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 1, -1, 0};
    return v[l];
}
__host__ __device__ constexpr int opp(int l) { return l == 18 ? 18 : (l ^ 1); }

// Lattice weight of direction l: 1/18 axis, 1/36 diagonal, 1/3 rest.
template <class T>
__host__ __device__ constexpr T weight(int l)
{
    return l == 18 ? (T)1 / (T)3 : ((l < 4 || l >= 16) ? (T)1 / (T)18 : (T)1 / (T)36);
}

// eq_dd_a0 / eq_dd_a1 / eq_dd4 / eq_dd5 / eq_dd18, common.h:132-160.
template <class T>
__host__ __device__ inline T eq_a0(T vela, T vela2, T rho_alpha)
{
    return ((T)1 / (T)18) * ((rho_alpha) + ((T)3) * (vela) + ((T)9 / (T)2) * (vela2));
}
template <class T>
__host__ __device__ inline T eq_a1(T vela, T vela2, T rho_alpha)
{
    return ((T)1 / (T)18) * ((rho_alpha) + ((T)-3) * (vela) + ((T)9 / (T)2) * (vela2));
}
template <class T>
__host__ __device__ inline T eq_4(T vab, T vab2, T rho_alpha)
{
    return ((T)1 / (T)36) * ((rho_alpha) + ((T)3) * (vab) + ((T)9 / (T)2) * (vab2));
}
template <class T>
__host__ __device__ inline T eq_5(T vab, T vab2, T rho_alpha)
{
    return ((T)1 / (T)36) * ((rho_alpha) + ((T)-3) * (vab) + ((T)9 / (T)2) * (vab2));
}
template <class T>
__host__ __device__ inline T eq_18(T rho_alpha)
{
    return ((T)1 / (T)3) * (rho_alpha);
}

// Density and momentum in the accumulation order of CLbmAlphaCPU.cpp:87-182
// (identical in CLbmBetaCPU.cpp:93-217).
template <class T>
__device__ __forceinline__ void moments(const T (&d)[kQ], T& rho, T& ux, T& uy, T& uz)
{
    T r, x, y, z;
    r = d[0];  x = d[0];
    r += d[1]; x -= d[1];
    r += d[2]; y = d[2];
    r += d[3]; y -= d[3];
    r += d[4]; x += d[4]; y += d[4];
    r += d[5]; x -= d[5]; y -= d[5];
    r += d[6]; x += d[6]; y -= d[6];
    r += d[7]; x -= d[7]; y += d[7];
    r += d[8]; x += d[8]; z = d[8];
    r += d[9]; x -= d[9]; z -= d[9];
    r += d[10]; x += d[10]; z -= d[10];
    r += d[11]; x -= d[11]; z += d[11];
    r += d[12]; y += d[12]; z += d[12];
    r += d[13]; y -= d[13]; z -= d[13];
    r += d[14]; y += d[14]; z -= d[14];
    r += d[15]; y -= d[15]; z += d[15];
    r += d[16]; z += d[16];
    r += d[17]; z -= d[17];
    r += d[18];
    rho = r; ux = x; uy = y; uz = z;
}

}  // namespace lbm
