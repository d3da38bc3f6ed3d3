// Per-cell device code of the A-A step shared by the single-domain kernels
// (lbm_step.hip) and the multi-cavity batch kernels (lbm_batch.hip):
// population accessors, one cell of alpha and of beta, fused x-face packing.
#pragma once

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
// Every population array is device (global) memory. The accessors say so
// explicitly: a base pointer read from memory (the batch member tables) is
// otherwise a generic pointer, and generic accesses compile to flat
// instructions, which wait on both memory counters.
#define LBM_GLOBAL __attribute__((address_space(1)))

template <class S, bool NT>
struct Acc32 {
    const char* const* plane;
    uint32_t off;
    __device__ __forceinline__ S load(int l, int32_t) const
    {
        const LBM_GLOBAL S* p = (const LBM_GLOBAL S*)(plane[l] + off);
        if (NT)
            return __builtin_nontemporal_load(p);
        return *p;
    }
    __device__ __forceinline__ void store(int l, int32_t, S v) const
    {
        // A fresh copy of the offset keeps the compiler from reusing the
        // loads' 64-bit addresses (19 VGPR pairs kept live across the
        // collision): each store then takes the uniform base + 32-bit offset
        // form, as the loads do.
        uint32_t o = off;
        asm volatile("" : "+v"(o));
        LBM_GLOBAL S* p = (LBM_GLOBAL S*)(const_cast<char*>(plane[l]) + o);
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
    __device__ __forceinline__ S load(int l, int32_t delta) const
    {
        return ((const LBM_GLOBAL S*)dd)[index(l, delta)];
    }
    __device__ __forceinline__ void store(int l, int32_t delta, S v) const
    {
        ((LBM_GLOBAL S*)dd)[index(l, delta)] = v;
    }
};

// ------------------------------------------------------- fused x packing ---

// Whether the cell lies in the LEFT (x = 1) / RIGHT (x = nx - 2) send plane
// with a fused pack (pl / pr) or unpack (ul / ur) buffer on that side;
// row = z * ny + y is its index in the dense face buffers. Returns pl || pr.
template <class C>
__device__ __forceinline__ bool xpack_where(const StepArgs<C>& a, uint32_t id, uint32_t& row,
                                            bool& pl, bool& pr, bool& ul, bool& ur)
{
    pl = pr = ul = ur = false;
    row = 0;
    if (!a.xpack_on)
        return false;
    row = a.div_nx.div(id);
    const int x = (int)(id - row * (uint32_t)a.nx);
    pl = x == 1 && a.xpack[0] != nullptr;
    pr = x == a.nx - 2 && a.xpack[1] != nullptr;
    ul = x == 1 && a.xunpack[0] != nullptr;
    ur = x == a.nx - 2 && a.xunpack[1] != nullptr;
    return pl || pr;
}

// After alpha: the slot values of a send-plane cell that the neighbour's halo
// needs, the planes whose vectors point back into this solver (c_x = +1 on the
// LEFT send plane, c_x = -1 on the RIGHT one), in the layout lbm_pack_face
// writes ([slot][row]); values pass through the storage type so 16-bit storage
// packs exactly what was stored.
template <class P>
__device__ __forceinline__ void xpack_alpha(const StepArgs<typename P::C>& a, uint32_t row,
                                            bool pl, bool pr, const typename P::C (&v)[kQ])
{
#pragma unroll
    for (int s = 0; s < kQ - 1; s++) {
        if (cx(s) == 0 || !(cx(s) > 0 ? pl : pr))
            continue;
        a.xpack[cx(s) > 0 ? 0 : 1][(int64_t)s * a.fcells + row] = P::load(P::store(v[s], s), s);
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

// ------------------------------------------------------ fused x unpacking ---

// The face work compiled into a step kernel: none, the fused x-face pack
// (lbm_set_step_pack), or the pack and the fused x-face unpack
// (lbm_set_step_unpack); a kernel without a part carries none of its code.
constexpr int kXfNone = 0, kXfPack = 1, kXfUnpack = 2;

// A received x-face buffer brings one plane of each opposite pair
// (xplane(k), xplane(k) + 1) = (c_x = +1, c_x = -1), k = 0..4: before alpha
// the send-plane cell's own slots of the vectors pointing into the solver
// through its face (c_x = +1 on the LEFT send plane, -1 on the RIGHT one:
// what lbm_unpack_face(phase 1) stores there); before beta the halo-plane
// slots the cell reads (c_x = -1 in the LEFT halo plane, +1 in the RIGHT one:
// what lbm_unpack_face(phase 0) stores there). `odd` selects the c_x = -1
// plane of each pair. Opposite vectors share their lattice weight, so both
// planes of a pair decode with xplane(k).
__host__ __device__ constexpr int xplane(int k) { return k == 0 ? 0 : 2 * k + 2; }

// The plane of each pair a face cell receives: before alpha (STEP 0) the
// LEFT face brings c_x = +1 and the RIGHT face c_x = -1, before beta the
// reverse.
template <int STEP>
__device__ __forceinline__ int xunpack_odd(bool ul)
{
    return STEP == 0 ? (ul ? 0 : 1) : (ul ? 1 : 0);
}

// Loads a face cell's 5 received values (ul: from the LEFT buffer, else the
// RIGHT one), issued with its slot loads. A beta cell's halo slots sit at the
// rows of their halo-plane cells (the reference linear wrap, as in
// xpack_beta).
template <class P, int STEP>
__device__ __forceinline__ void xunpack_load(const StepArgs<typename P::C>& a, uint32_t row,
                                             bool ul, int odd, typename P::C (&u)[5])
{
    const typename P::C* buf = a.xunpack[ul ? 0 : 1];
#pragma unroll
    for (int k = 0; k < 5; k++) {
        const int e = xplane(k);
        int32_t t = (int32_t)row;
        if (STEP == 1) {
            const int32_t off = cy(e) + cz(e) * a.ny;   // the c_x = -1 plane's is -off
            t += odd ? -off : off;
            t = t < 0 ? t + a.fcells : t;
            t = t >= a.fcells ? t - a.fcells : t;
        }
        u[k] = buf[(int64_t)(e + odd) * a.fcells + t];
    }
}

// The value a step would load from the slot after lbm_unpack_face: the
// received one through the storage type, its range marked in ov as the
// unpack marks it.
template <class P>
__device__ __forceinline__ typename P::C xunpack_value(typename P::C r, int l, uint32_t& ov)
{
    return P::load(P::store_checked(r, l, ov), l);
}

// ----------------------------------------------------------------- alpha ---

// One cell of CLbmAlphaCPU::alphaKernelCPU (CLbmAlphaCPU.cpp:75-494). The
// reference's `tmp` is a #define for `rho` (:199): every gravity term
// overwrites rho and the stored density is the last gravity term. `flag` is
// the cell's flag, passed in so a caller that already holds it skips a load.
template <class P, int XF = kXfPack, class A>
__device__ __forceinline__ void alpha_cell(const StepArgs<typename P::C>& a, uint32_t id,
                                           int flag, const A& acc)
{
    using C = typename P::C;
    uint32_t row = 0;
    bool pl = false, pr = false, ul = false, ur = false;
    if (XF != kXfNone)
        xpack_where(a, id, row, pl, pr, ul, ur);
    if (XF != kXfUnpack)
        ul = ur = false;
    const bool pack = pl || pr;
    const bool unpack = ul || ur;
    const int odd = xunpack_odd<0>(ul);
    // The slots are loaded for every cell, whatever its flag, so that the
    // flag's latency overlaps theirs instead of preceding them. GHOST_LAYER
    // (skipped by the reference, :75-76) and OBSTACLE without outputs then
    // store nothing below.
    C d[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        d[l] = P::load(acc.load(l, 0), l);
    C u[5];
    if (unpack)
        xunpack_load<P, 0>(a, row, ul, odd, u);
    // keeps the flag test (and its wait) below the issued loads
    __builtin_amdgcn_sched_barrier(0);
    if (unpack) {
        uint32_t uov = 0;
        // a cell that stores nothing (GHOST_LAYER, OBSTACLE) leaves the
        // received values where the unpack would have put them: the unpack's
        // codes, encoded from the buffer values (decoding and encoding again
        // need not give them back); written here, so that the received values
        // are not kept across the collision
        const bool keep = flag == kGhostLayer || flag == kObstacle;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int e = xplane(k);
            const C v = xunpack_value<P>(u[k], e, uov);
            d[e] = odd ? d[e] : v;
            d[e + 1] = odd ? v : d[e + 1];
            if (keep) {
                if (odd)
                    acc.store(e + 1, 0, P::store(u[k], e));
                else
                    acc.store(e, 0, P::store(u[k], e));
            }
        }
        raise_overflow<P>(a.status, uov);
    }
    if (flag == kGhostLayer) {   // packs its unchanged slots on a send plane
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
        uint32_t ov = 0;
#pragma unroll
        for (int l = 0; l < kQ; l++)
            acc.store(opp(l), 0, P::store_checked(e[l], opp(l), ov));
        raise_overflow<P>(a.status, ov);
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


// ------------------------------------------------------------------ beta ---

// One cell of CLbmBetaCPU::betaKernelCPU (CLbmBetaCPU.cpp:83-451).
// `dd_param` is a #define for `rho` (:224): the stored FLUID density is
// rho - 1.5 |j|^2. OBSTACLE swaps opposite pairs, which writes back exactly
// what was read, so its stores are skipped. GHOST_LAYER passes values through.
// Neighbour indices wrap modulo the cell count (domainWrap, :44-50). `flag`
// is the cell's flag, as for alpha_cell.
template <class P, int XF = kXfPack, class A>
__device__ __forceinline__ void beta_cell(const StepArgs<typename P::C>& a, uint32_t id,
                                          int flag, const A& acc)
{
    using C = typename P::C;
    uint32_t row = 0;
    bool pl = false, pr = false, ul = false, ur = false;
    if (XF != kXfNone)
        xpack_where(a, id, row, pl, pr, ul, ur);
    if (XF != kXfUnpack)
        ul = ur = false;
    const bool pack = pl || pr;
    const bool unpack = ul || ur;
    const int odd = xunpack_odd<1>(ul);
    // plane l at id + c_l holds direction opp(l) arriving at this cell. As in
    // alpha_cell the slots are loaded whatever the flag (the caller's accessor
    // keeps every neighbour index in bounds); OBSTACLE, which writes back what
    // it reads, stores nothing below.
    C d[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        d[opp(l)] = P::load(acc.load(l, cx(l) + cy(l) * a.nx + cz(l) * a.nxy), l);
    C u[5];
    if (unpack)
        xunpack_load<P, 1>(a, row, ul, odd, u);
    __builtin_amdgcn_sched_barrier(0);
    if (unpack) {
        // slot l's value is d[opp(l)]: slot e in d[e + 1], slot e + 1 in d[e].
        // OBSTACLE stores nothing: the received values stay where the unpack
        // puts them (written here, so they are not kept across the collision)
        uint32_t uov = 0;
        const bool keep = flag == kObstacle;
#pragma unroll
        for (int k = 0; k < 5; k++) {
            const int e = xplane(k);
            const C v = xunpack_value<P>(u[k], e, uov);
            d[e + 1] = odd ? d[e + 1] : v;
            d[e] = odd ? v : d[e];
            if (keep) {
                if (odd)
                    acc.store(e + 1, cx(e + 1) + cy(e + 1) * a.nx + cz(e + 1) * a.nxy,
                              P::store(u[k], e));
                else
                    acc.store(e, cx(e) + cy(e) * a.nx + cz(e) * a.nxy, P::store(u[k], e));
            }
        }
        raise_overflow<P>(a.status, uov);
    }

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
        uint32_t ov = 0;
#pragma unroll
        for (int l = 0; l < kQ; l++)
            acc.store(l, cx(l) + cy(l) * a.nx + cz(l) * a.nxy, P::store_checked(d[l], l, ov));
        raise_overflow<P>(a.status, ov);
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


}  // namespace lbm
