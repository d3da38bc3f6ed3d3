// Fused collide-stream A-A kernels (init, alpha, beta) for gfx950.
//
// Layout: structure of arrays, plane l of the populations at
// dd + l * stride, cell id = (z * ny_h + y) * nx_h + x (x fastest), one byte
// of flag per cell. One thread owns one cell; a 64-lane wavefront covers 64
// consecutive cells, so every population plane is read and written as
// contiguous 256-byte runs (4-byte elements).
//
// Alpha (even step, CLbmAlphaCPU.cpp:40-502): the cell reads its own 19 slots
// and writes the post-collision value of direction l into slot opp(l).
// Beta (odd step, CLbmBetaCPU.cpp:52-459): the cell reads plane l at
// id + c_l (direction opp(l) arriving) and, after collision, writes direction
// l into the same address. So in both kernels a thread reads and writes the
// same 19 addresses and no two threads share an address: the step is
// race-free in place without any barrier.
//
// Both steps are pure HBM streams: 19 loads + 19 stores per cell
// (152 B/cell in FP32), ~1.6 flop/B, no matrix-core work.
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"
#include "lbm_storage.hpp"

namespace lbm {

constexpr int kBlock = 256;

// Workgroups are dispatched round-robin over the 8 XCDs (each with its own
// L2). Renumber them so every XCD works on one contiguous range of cells:
// neighbouring cells then share partial cache lines inside one L2.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

template <class C>
__device__ __forceinline__ uint32_t box_cell(const StepArgs<C>& a, uint32_t gid)
{
    const uint32_t t = a.div_bx.div(gid);
    const uint32_t x = gid - t * a.div_bx.d;
    const uint32_t z = a.div_by.div(t);
    const uint32_t y = t - z * a.div_by.d;
    return (uint32_t)((a.oz + (int32_t)z) * a.nxy + (a.oy + (int32_t)y) * a.nx + a.ox + (int32_t)x);
}

// Population accessors. Acc32 keeps one uniform base pointer per plane (the
// beta neighbour shift is folded into it) and one 32-bit byte offset per lane
// shared by all 19 planes, so every access is a saddr global load/store with
// no per-plane vector address arithmetic; it needs cells * sizeof(S) < 2^32.
// Acc64 uses a 64-bit element index per plane and, with WRAP, the reference's
// linear-index wrap (domainWrap, CLbmBetaCPU.cpp:44-50).
// The plane bases come precomputed in the kernel arguments: opaque uniform
// values the compiler cannot fold into per-lane pointer chains.
template <class S, bool NT>
struct Acc32 {
    const char* const* plane;
    uint32_t off;
    __device__ __forceinline__ S load(int l, int32_t) const
    {
        const S* p = (const S*)(plane[l] + off);
        if (NT)
            return __builtin_nontemporal_load(p);
        return *p;
    }
    __device__ __forceinline__ void store(int l, int32_t, S v) const
    {
        S* p = (S*)(const_cast<char*>(plane[l]) + off);
        if (NT)
            __builtin_nontemporal_store(v, p);
        else
            *p = v;
    }
};

template <class S, bool WRAP>
struct Acc64 {
    S* dd;
    int64_t stride;
    int32_t id, n;
    __device__ __forceinline__ int64_t index(int l, int32_t delta) const
    {
        int32_t j = id + delta;
        if (WRAP) {
            j = (j >= n) ? j - n : j;
            j = (j < 0) ? j + n : j;
        }
        return (int64_t)l * stride + j;
    }
    __device__ __forceinline__ S load(int l, int32_t delta) const { return dd[index(l, delta)]; }
    __device__ __forceinline__ void store(int l, int32_t delta, S v) const { dd[index(l, delta)] = v; }
};

// ------------------------------------------------------------------ init ---

// lbm_init.cu:41-208 / CLbmInitCPU.cpp:37-171: flags from the six face
// conditions in x-fastest order (the first OBSTACLE wins, GHOST_LAYER can be
// overwritten), populations at the rest equilibrium computed with the
// reference's equilibrium arithmetic, u = 0, rho = 1.
template <class P>
__global__ __launch_bounds__(kBlock) void k_init(StepArgs<typename P::C> a, int ny,
                                                 int nz, int4 bc_lo, int2 bc_hi)
{
    using C = typename P::C;
    using S = typename P::S;
    const uint32_t id = blockIdx.x * kBlock + threadIdx.x;
    if (id >= (uint32_t)a.n)
        return;
    const int x = (int)(id % (uint32_t)a.nx);
    const int y = (int)((id / (uint32_t)a.nx) % (uint32_t)ny);
    const int z = (int)(id / (uint32_t)a.nxy);
    const int bc[6] = {bc_lo.x, bc_lo.y, bc_lo.z, bc_lo.w, bc_hi.x, bc_hi.y};
    int flag = kFluid;
    if (x == 0)
        flag = bc[0];
    if (x == a.nx - 1 && flag != kObstacle)
        flag = bc[1];
    if (y == 0 && flag != kObstacle)
        flag = bc[2];
    if (y == ny - 1 && flag != kObstacle)
        flag = bc[3];
    if (z == 0 && flag != kObstacle)
        flag = bc[4];
    if (z == nz - 1 && flag != kObstacle)
        flag = bc[5];
    const_cast<uint8_t*>(a.flags)[id] = (uint8_t)flag;

    S* dd = (S*)a.dd;
    const C ux = (C)0, uy = (C)0, uz = (C)0, rho = (C)1;
    C vela2 = ux * ux;
    const C dd_param = rho - ((C)3 / (C)2) * (vela2);
    C f[kQ];
    f[0] = eq_a0(ux, vela2, dd_param);
    f[1] = eq_a1(ux, vela2, dd_param);
    vela2 = uy * uy;
    f[2] = eq_a0(uy, vela2, dd_param);
    f[3] = eq_a1(uy, vela2, dd_param);
    C vab = ux + uy, vab2 = vab * vab;
    f[4] = eq_4(vab, vab2, dd_param);
    f[5] = eq_5(vab, vab2, dd_param);
    vab = ux - uy; vab2 = vab * vab;
    f[6] = eq_4(vab, vab2, dd_param);
    f[7] = eq_5(vab, vab2, dd_param);
    vab = ux + uz; vab2 = vab * vab;
    f[8] = eq_4(vab, vab2, dd_param);
    f[9] = eq_5(vab, vab2, dd_param);
    vab = ux - uz; vab2 = vab * vab;
    f[10] = eq_4(vab, vab2, dd_param);
    f[11] = eq_5(vab, vab2, dd_param);
    vab = uy + uz; vab2 = vab * vab;
    f[12] = eq_4(vab, vab2, dd_param);
    f[13] = eq_5(vab, vab2, dd_param);
    vab = uy - uz; vab2 = vab * vab;
    f[14] = eq_4(vab, vab2, dd_param);
    f[15] = eq_5(vab, vab2, dd_param);
    vela2 = uz * uz;
    f[16] = eq_a0(uz, vela2, dd_param);
    f[17] = eq_a1(uz, vela2, dd_param);
    f[18] = eq_18(dd_param);
#pragma unroll
    for (int l = 0; l < kQ; l++)
        dd[(int64_t)l * a.stride + id] = P::store(f[l], l);
    if (a.store_u) {
        a.u[id] = ux;
        a.u[a.stride + id] = uy;
        a.u[2 * a.stride + id] = uz;
    }
    if (a.store_rho)
        a.rho[id] = rho;
}

// ------------------------------------------------------- fused x packing ---

// Whether the cell lies in the LEFT (x = 1) / RIGHT (x = nx - 2) send plane;
// row = z * ny + y is its index in the dense face buffers.
template <class C>
__device__ __forceinline__ bool xpack_where(const StepArgs<C>& a, uint32_t id, uint32_t& row,
                                            bool& pl, bool& pr)
{
    pl = pr = false;
    row = 0;
    if (!a.xpack_on)
        return false;
    row = a.div_nx.div(id);
    const int x = (int)(id - row * (uint32_t)a.nx);
    pl = x == 1 && a.xpack[0] != nullptr;
    pr = x == a.nx - 2 && a.xpack[1] != nullptr;
    return pl || pr;
}

// After alpha: the 19 slot values of a send-plane cell, in the layout
// lbm_pack_face writes ([slot][row]); values pass through the storage type so
// 16-bit storage packs exactly what was stored.
template <class P>
__device__ __forceinline__ void xpack_alpha(const StepArgs<typename P::C>& a, uint32_t row,
                                            bool pl, bool pr, const typename P::C (&v)[kQ])
{
    using C = typename P::C;
#pragma unroll
    for (int s = 0; s < kQ; s++) {
        const C w = P::load(P::store(v[s], s), s);
        if (pl)
            a.xpack[0][(int64_t)s * a.fcells + row] = w;
        if (pr)
            a.xpack[1][(int64_t)s * a.fcells + row] = w;
    }
}

// After beta: the populations that cross the x faces. From a LEFT send-plane
// cell the directions with c_x = -1 land in the LEFT halo plane, from a RIGHT
// one those with c_x = +1 land in the RIGHT halo plane; each is written at
// its landing cell's row (reference linear wrap), i.e. exactly the entries
// of the halo-plane pack that the neighbour's direction-filtered unpack reads.
template <class P>
__device__ __forceinline__ void xpack_beta(const StepArgs<typename P::C>& a, uint32_t row,
                                           bool pl, bool pr, const typename P::C (&v)[kQ])
{
    using C = typename P::C;
#pragma unroll
    for (int l = 0; l < kQ - 1; l++) {
        const bool side = cx(l) < 0 ? pl : (cx(l) > 0 ? pr : false);
        if (cx(l) != 0 && side) {
            int32_t t = (int32_t)row + cy(l) + cz(l) * a.ny;
            t = t < 0 ? t + a.fcells : t;
            t = t >= a.fcells ? t - a.fcells : t;
            a.xpack[cx(l) < 0 ? 0 : 1][(int64_t)l * a.fcells + t] = P::load(P::store(v[l], l), l);
        }
    }
}

// ----------------------------------------------------------------- alpha ---

// One cell of CLbmAlphaCPU::alphaKernelCPU (CLbmAlphaCPU.cpp:75-494). The
// reference's `tmp` is a #define for `rho` (:199): every gravity term
// overwrites rho and the stored density is the last gravity term.
template <class P, class A>
__device__ __forceinline__ void alpha_cell(const StepArgs<typename P::C>& a, uint32_t id,
                                           const A& acc)
{
    using C = typename P::C;
    const int flag = a.flags[id];
    uint32_t row;
    bool pl, pr;
    const bool pack = xpack_where(a, id, row, pl, pr);
    // GHOST_LAYER is skipped (:75-76); OBSTACLE without outputs does nothing
    if (!pack && (flag == kGhostLayer || (flag == kObstacle && !a.store_u && !a.store_rho)))
        return;

    C d[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        d[l] = P::load(acc.load(l, 0), l);
    if (flag == kGhostLayer) {   // reached only to pack its unchanged slots
        xpack_alpha<P>(a, row, pl, pr, d);
        return;
    }
    C rho, ux, uy, uz;
    moments(d, rho, ux, uy, uz);

    if (flag == kFluid || flag == kVelocityInjection) {
        const C inv_tau = a.inv_tau;
        const C g0 = a.g0, g1 = a.g1, g2 = a.g2;
        const bool fluid = (flag == kFluid);
        C vel2, vela2, vab, dd_param;
        C e[kQ];
        if (!fluid) {
            ux = a.lid;
            uy = (C)0;
            uz = (C)0;
            rho = (C)1;
        }
        vel2 = ux * ux + uy * uy + uz * uz;
        dd_param = rho - ((C)3 / (C)2) * vel2;
        // post-collision value of direction l: relaxation for FLUID (d += inv_tau*(eq - d)),
        // the equilibrium itself for VELOCITY_INJECTION
#define LBM_RELAX(l, eqv) e[l] = fluid ? (d[l] + inv_tau * ((eqv) - d[l])) : (eqv)
        rho = g0 * ((C)1 / (C)18) * rho;
        vela2 = ux * ux;
        LBM_RELAX(1, eq_a1(ux, vela2, dd_param));
        e[1] -= rho;
        LBM_RELAX(0, eq_a0(ux, vela2, dd_param));
        e[0] += rho;

        rho = g1 * ((C)-1 / (C)18) * rho;
        vela2 = uy * uy;
        LBM_RELAX(3, eq_a1(uy, vela2, dd_param));
        e[3] -= rho;
        LBM_RELAX(2, eq_a0(uy, vela2, dd_param));
        e[2] += rho;

        vab = ux + uy;
        vela2 = vab * vab;
        rho = (g0 - g1) * ((C)1 / (C)36) * rho;
        LBM_RELAX(5, eq_5(vab, vela2, dd_param));
        e[5] -= rho;
        LBM_RELAX(4, eq_4(vab, vela2, dd_param));
        e[4] += rho;

        vab = ux - uy;
        vela2 = vab * vab;
        rho = (g0 + g1) * ((C)1 / (C)36) * rho;
        LBM_RELAX(7, eq_5(vab, vela2, dd_param));
        e[7] -= rho;
        LBM_RELAX(6, eq_4(vab, vela2, dd_param));
        e[6] += rho;

        vab = ux + uz;
        vela2 = vab * vab;
        rho = (g0 + g2) * ((C)1 / (C)36) * rho;
        LBM_RELAX(9, eq_5(vab, vela2, dd_param));
        e[9] -= rho;
        LBM_RELAX(8, eq_4(vab, vela2, dd_param));
        e[8] += rho;

        rho = (g0 - g2) * ((C)1 / (C)36) * rho;
        vab = ux - uz;
        vela2 = vab * vab;
        LBM_RELAX(11, eq_5(vab, vela2, dd_param));
        e[11] -= rho;
        LBM_RELAX(10, eq_4(vab, vela2, dd_param));
        e[10] += rho;

        vab = uy + uz;
        vela2 = vab * vab;
        rho = (g2 - g1) * ((C)1 / (C)36) * rho;
        LBM_RELAX(13, eq_5(vab, vela2, dd_param));
        e[13] -= rho;
        LBM_RELAX(12, eq_4(vab, vela2, dd_param));
        e[12] += rho;

        vab = uy - uz;
        vela2 = vab * vab;
        rho = (g2 + g1) * ((C)-1 / (C)36) * rho;
        LBM_RELAX(15, eq_5(vab, vela2, dd_param));
        e[15] -= rho;
        LBM_RELAX(14, eq_4(vab, vela2, dd_param));
        e[14] += rho;

        vela2 = uz * uz;
        rho = g2 * ((C)1 / (C)18) * rho;
        LBM_RELAX(17, eq_a1(uz, vela2, dd_param));
        e[17] -= rho;
        LBM_RELAX(16, eq_a0(uz, vela2, dd_param));
        e[16] += rho;

        LBM_RELAX(18, eq_18(dd_param));
#undef LBM_RELAX
        // direction l goes to slot opp(l) of this cell
#pragma unroll
        for (int l = 0; l < kQ; l++)
            acc.store(opp(l), 0, P::store(e[l], opp(l)));
        if (pack) {
            C out[kQ];
#pragma unroll
            for (int s2 = 0; s2 < kQ; s2++)
                out[s2] = e[opp(s2)];
            xpack_alpha<P>(a, row, pl, pr, out);
        }
    } else if (flag == kObstacle) {
        // no stores: the A-A layout makes bounce-back implicit (:330-343)
        if (pack)
            xpack_alpha<P>(a, row, pl, pr, d);
        if (a.store_u) {
            ux = (C)0;
            uy = (C)0;
            uz = (C)0;
        }
    }

    if (a.store_u) {
        a.u[id] = ux;
        a.u[a.stride + id] = uy;
        a.u[2 * a.stride + id] = uz;
    }
    if (a.store_rho)
        a.rho[id] = rho;
}

template <class P, bool FULL, bool OFF32, bool NT>
__global__ __launch_bounds__(kBlock) void k_alpha(StepArgs<typename P::C> a)
{
    using S = typename P::S;
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t gid = blk * kBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    const uint32_t id = FULL ? gid : box_cell(a, gid);
    if (OFF32) {
        const Acc32<S, NT> acc{a.plane, id * (uint32_t)sizeof(S)};
        alpha_cell<P>(a, id, acc);
    } else {
        const Acc64<S, false> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
        alpha_cell<P>(a, id, acc);
    }
}

// ------------------------------------------------------------------ beta ---

// One cell of CLbmBetaCPU::betaKernelCPU (CLbmBetaCPU.cpp:83-451).
// `dd_param` is a #define for `rho` (:224): the stored FLUID density is
// rho - 1.5 |j|^2. OBSTACLE swaps opposite pairs, which writes back exactly
// what was read, so its stores are skipped. GHOST_LAYER passes values through.
// Neighbour indices wrap modulo the cell count (domainWrap, :44-50).
template <class P, class A>
__device__ __forceinline__ void beta_cell(const StepArgs<typename P::C>& a, uint32_t id,
                                          const A& acc)
{
    using C = typename P::C;
    const int flag = a.flags[id];
    uint32_t row;
    bool pl, pr;
    const bool pack = xpack_where(a, id, row, pl, pr);
    // OBSTACLE writes back what it reads; without outputs it does nothing
    if (!pack && flag == kObstacle && !a.store_u && !a.store_rho)
        return;

    // plane l at id + c_l holds direction opp(l) arriving at this cell
    C d[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        d[opp(l)] = P::load(acc.load(l, cx(l) + cy(l) * a.nx + cz(l) * a.nxy), l);

    C rho, ux, uy, uz;
    moments(d, rho, ux, uy, uz);

    if (flag == kFluid) {
        const C inv_tau = a.inv_tau;
        C vel2 = ux * ux + uy * uy + uz * uz;
        rho = rho - ((C)3 / (C)2) * (vel2);
        C vela2 = ux * ux;
        d[0] += inv_tau * (eq_a0(ux, vela2, rho) - d[0]);
        d[1] += inv_tau * (eq_a1(ux, vela2, rho) - d[1]);
        vela2 = uy * uy;
        d[2] += inv_tau * (eq_a0(uy, vela2, rho) - d[2]);
        d[3] += inv_tau * (eq_a1(uy, vela2, rho) - d[3]);
        vel2 = ux + uy;
        vela2 = vel2 * vel2;
        d[4] += inv_tau * (eq_4(vel2, vela2, rho) - d[4]);
        d[5] += inv_tau * (eq_5(vel2, vela2, rho) - d[5]);
        vel2 = ux - uy;
        vela2 = vel2 * vel2;
        d[6] += inv_tau * (eq_4(vel2, vela2, rho) - d[6]);
        d[7] += inv_tau * (eq_5(vel2, vela2, rho) - d[7]);
        vel2 = ux + uz;
        vela2 = vel2 * vel2;
        d[8] += inv_tau * (eq_4(vel2, vela2, rho) - d[8]);
        d[9] += inv_tau * (eq_5(vel2, vela2, rho) - d[9]);
        vel2 = ux - uz;
        vela2 = vel2 * vel2;
        d[10] += inv_tau * (eq_4(vel2, vela2, rho) - d[10]);
        d[11] += inv_tau * (eq_5(vel2, vela2, rho) - d[11]);
        vel2 = uy + uz;
        vela2 = vel2 * vel2;
        d[12] += inv_tau * (eq_4(vel2, vela2, rho) - d[12]);
        d[13] += inv_tau * (eq_5(vel2, vela2, rho) - d[13]);
        vel2 = uy - uz;
        vela2 = vel2 * vel2;
        d[14] += inv_tau * (eq_4(vel2, vela2, rho) - d[14]);
        d[15] += inv_tau * (eq_5(vel2, vela2, rho) - d[15]);
        vela2 = uz * uz;
        d[16] += inv_tau * (eq_a0(uz, vela2, rho) - d[16]);
        d[17] += inv_tau * (eq_a1(uz, vela2, rho) - d[17]);
        d[18] += inv_tau * (eq_18(rho) - d[18]);
    } else if (flag == kVelocityInjection) {
        ux = a.lid;
        uy = (C)0;
        uz = (C)0;
        rho = (C)1;
        C vel2 = ux * ux + uy * uy + uz * uz;
        rho = rho - ((C)3 / (C)2) * vel2;
        C vela2 = ux * ux;
        d[0] = eq_a0(ux, vela2, rho);
        d[1] = eq_a1(ux, vela2, rho);
        vela2 = uy * uy;
        d[2] = eq_a0(uy, vela2, rho);
        d[3] = eq_a1(uy, vela2, rho);
        vel2 = ux + uy;
        vela2 = vel2 * vel2;
        d[4] = eq_4(vel2, vela2, rho);
        d[5] = eq_5(vel2, vela2, rho);
        vel2 = ux - uy;
        vela2 = vel2 * vel2;
        d[6] = eq_4(vel2, vela2, rho);
        d[7] = eq_5(vel2, vela2, rho);
        vel2 = ux + uz;
        vela2 = vel2 * vel2;
        d[8] = eq_4(vel2, vela2, rho);
        d[9] = eq_5(vel2, vela2, rho);
        vel2 = ux - uz;
        vela2 = vel2 * vel2;
        d[10] = eq_4(vel2, vela2, rho);
        d[11] = eq_5(vel2, vela2, rho);
        vel2 = uy + uz;
        vela2 = vel2 * vel2;
        d[12] = eq_4(vel2, vela2, rho);
        d[13] = eq_5(vel2, vela2, rho);
        vel2 = uy - uz;
        vela2 = vel2 * vel2;
        d[14] = eq_4(vel2, vela2, rho);
        d[15] = eq_5(vel2, vela2, rho);
        vela2 = uz * uz;
        d[16] = eq_a0(uz, vela2, rho);
        d[17] = eq_a1(uz, vela2, rho);
        d[18] = eq_18(rho);
    } else if (flag == kObstacle) {
        if (pack) {
            // the swap writes back the value read from plane l (:302-311)
            C out[kQ];
#pragma unroll
            for (int l = 0; l < kQ; l++)
                out[l] = d[opp(l)];
            xpack_beta<P>(a, row, pl, pr, out);
        }
        if (a.store_u) {
            ux = (C)0;
            uy = (C)0;
            uz = (C)0;
        }
    }

    if (flag != kObstacle) {
        // direction l goes to slot l of neighbour id + c_l
#pragma unroll
        for (int l = 0; l < kQ; l++)
            acc.store(l, cx(l) + cy(l) * a.nx + cz(l) * a.nxy, P::store(d[l], l));
        if (pack)
            xpack_beta<P>(a, row, pl, pr, d);
    }

    if (a.store_u) {
        a.u[id] = ux;
        a.u[a.stride + id] = uy;
        a.u[2 * a.stride + id] = uz;
    }
    if (a.store_rho)
        a.rho[id] = rho;
}

template <class P, bool FULL, bool OFF32, bool NT>
__global__ __launch_bounds__(kBlock) void k_beta(StepArgs<typename P::C> a)
{
    using S = typename P::S;
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t gid = blk * kBlock + threadIdx.x;
    // One decision per workgroup: neighbours of every cell of this block stay
    // inside [0, n) unless the block touches the first or last slab.
    const uint32_t g_first = blk * kBlock;
    const uint32_t g_last = min(g_first + kBlock - 1, a.cells - 1);
    const uint32_t id_first = FULL ? g_first : box_cell(a, g_first);
    const uint32_t id_last = FULL ? g_last : box_cell(a, g_last);
    const int64_t reach = (int64_t)a.nxy + a.nx + 1;
    const bool wrap = (int64_t)id_first < reach || (int64_t)id_last + reach >= (int64_t)a.n;
    if (gid >= a.cells)
        return;
    const uint32_t id = FULL ? gid : box_cell(a, gid);
    if (wrap) {
        const Acc64<S, true> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
        beta_cell<P>(a, id, acc);
    } else if (OFF32) {
        const Acc32<S, NT> acc{a.plane, id * (uint32_t)sizeof(S)};
        beta_cell<P>(a, id, acc);
    } else {
        const Acc64<S, false> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
        beta_cell<P>(a, id, acc);
    }
}

// -------------------------------------------------------------- launches ---

template <class P>
static StepArgs<typename P::C> make_args(const lbm_solver* s, const Box& b)
{
    using C = typename P::C;
    StepArgs<C> a;
    a.dd = s->dd;
    a.flags = s->flags;
    a.rho = (C*)s->rho;
    a.u = (C*)s->u;
    a.stride = s->stride;
    a.n = (int32_t)s->n;
    a.nx = s->size_h[0];
    a.nxy = s->size_h[0] * s->size_h[1];
    a.ox = b.ox;
    a.oy = b.oy;
    a.oz = b.oz;
    a.cells = b.cells();
    a.div_bx = FastDiv((uint32_t)b.bx);
    a.div_by = FastDiv((uint32_t)b.by);
    a.inv_tau = (C)s->p.inv_tau;
    a.lid = (C)s->p.lid_velocity;
    a.g0 = (C)s->p.gravitation[0];
    a.g1 = (C)s->p.gravitation[1];
    a.g2 = (C)s->p.gravitation[2];
    a.store_rho = s->p.store_densities;
    a.store_u = s->p.store_velocities;
    a.xcd = 0;
    a.xpack[0] = (C*)s->xpack[0];
    a.xpack[1] = (C*)s->xpack[1];
    a.xpack_on = (s->xpack[0] || s->xpack[1]) ? 1 : 0;
    a.div_nx = FastDiv((uint32_t)s->size_h[0]);
    a.ny = s->size_h[1];
    a.fcells = s->size_h[1] * s->size_h[2];
    for (int l = 0; l < kQ; l++)
        a.plane[l] = (const char*)s->dd + (int64_t)l * s->stride * (int64_t)sizeof(typename P::S);
    return a;
}

static bool is_full(const lbm_solver* s, const Box& b)
{
    return b.ox == 0 && b.oy == 0 && b.oz == 0 && b.bx == s->size_h[0] &&
           b.by == s->size_h[1] && b.bz == s->size_h[2];
}

template <class P>
static hipError_t init_t(const lbm_solver* s, hipStream_t st)
{
    Box all{0, 0, 0, s->size_h[0], s->size_h[1], s->size_h[2]};
    auto a = make_args<P>(s, all);
    const uint32_t blocks = (uint32_t)((s->n + kBlock - 1) / kBlock);
    int4 lo = make_int4(s->p.bc[0], s->p.bc[1], s->p.bc[2], s->p.bc[3]);
    int2 hi = make_int2(s->p.bc[4], s->p.bc[5]);
    hipLaunchKernelGGL(k_init<P>, dim3(blocks), dim3(kBlock), 0, st, a, s->size_h[1],
                       s->size_h[2], lo, hi);
    return hipGetLastError();
}

// 32-bit byte offsets reach every element of a plane
static bool off32(const lbm_solver* s)
{
    return (uint64_t)s->stride * (uint64_t)s->elem_bytes < (1ull << 32);
}

template <class P, template <class, bool, bool, bool> class K, bool FULL>
static void launch_variant(const lbm_solver* s, const StepArgs<typename P::C>& a, int opts,
                           hipStream_t st)
{
    const uint32_t blocks = (a.cells + kBlock - 1) / kBlock;
    if (!off32(s))
        hipLaunchKernelGGL((K<P, FULL, false, false>::kernel), dim3(blocks), dim3(kBlock), 0,
                           st, a);
    else if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((K<P, FULL, true, true>::kernel), dim3(blocks), dim3(kBlock), 0, st,
                           a);
    else
        hipLaunchKernelGGL((K<P, FULL, true, false>::kernel), dim3(blocks), dim3(kBlock), 0,
                           st, a);
}

template <class P, bool FULL, bool OFF32, bool NT>
struct AlphaK {
    static constexpr auto kernel = k_alpha<P, FULL, OFF32, NT>;
};
template <class P, bool FULL, bool OFF32, bool NT>
struct BetaK {
    static constexpr auto kernel = k_beta<P, FULL, OFF32, NT>;
};

template <class P>
static hipError_t alpha_t(const lbm_solver* s, const Box& b, hipStream_t st)
{
    auto a = make_args<P>(s, b);
    a.xcd = (s->launch[0] & LBM_LAUNCH_XCD_REMAP) ? 1 : 0;
    if (is_full(s, b))
        launch_variant<P, AlphaK, true>(s, a, s->launch[0], st);
    else
        launch_variant<P, AlphaK, false>(s, a, s->launch[0], st);
    return hipGetLastError();
}

template <class P>
static hipError_t beta_t(const lbm_solver* s, const Box& b, hipStream_t st)
{
    auto a = make_args<P>(s, b);
    a.xcd = (s->launch[1] & LBM_LAUNCH_XCD_REMAP) ? 1 : 0;
    // beta reads and writes plane l at id + c_l: fold the shift into the base
    for (int l = 0; l < kQ; l++)
        a.plane[l] += (int64_t)(cx(l) + cy(l) * a.nx + cz(l) * a.nxy) *
                      (int64_t)sizeof(typename P::S);
    if (is_full(s, b))
        launch_variant<P, BetaK, true>(s, a, s->launch[1], st);
    else
        launch_variant<P, BetaK, false>(s, a, s->launch[1], st);
    return hipGetLastError();
}

#define LBM_DISPATCH(fn, ...)                                  \
    switch (s->p.precision) {                                  \
    case LBM_F32: return fn<StoreF32>(__VA_ARGS__);            \
    case LBM_F64: return fn<StoreF64>(__VA_ARGS__);            \
    case LBM_F16S: return fn<StoreF16S>(__VA_ARGS__);          \
    case LBM_BF16S: return fn<StoreBF16S>(__VA_ARGS__);        \
    default: return hipErrorInvalidValue;                      \
    }

hipError_t launch_init(const lbm_solver* s, hipStream_t st) { LBM_DISPATCH(init_t, s, st) }
hipError_t launch_alpha(const lbm_solver* s, const Box& b, hipStream_t st)
{
    LBM_DISPATCH(alpha_t, s, b, st)
}
hipError_t launch_beta(const lbm_solver* s, const Box& b, hipStream_t st)
{
    LBM_DISPATCH(beta_t, s, b, st)
}

}  // namespace lbm
