// Two time steps (alpha, then beta) of a whole cavity in one pass over HBM,
// out of place: populations are read from one buffer and the state after the
// beta half-step is written to a second buffer, in the same natural layout
// the in-place A-A pair leaves behind. Used by the batch launch for members
// that own their whole domain (no face exchange between the half-steps).
//
// A workgroup owns a column tile of TX x TY cells (PairShape) and a chunk of
// z-planes, and sweeps it in z. Each thread runs alpha_cell on one cell of
// the window (the tile plus a one-cell ring, kWin cells) of the next plane,
// from populations it fetched into registers one plane ahead, and writes the
// post-alpha slots to LDS. beta_cell of the tile cells of the current plane
// then gathers its 19 inputs from LDS and writes its outputs straight to the
// destination buffer. Beta of plane z reads slot l of plane z + c_l.z only,
// so the slots are kept per group of c_l.z: the 5 slots with c_l.z = +1 for
// one plane, the 9 with c_l.z = 0 for two, the 5 with c_l.z = -1 for three
// (38 slot-planes of LDS instead of 57 for three whole planes).
//
// Every (cell y, slot l) of the destination is written exactly once:
//  * by beta of x = y - c_l when x is not OBSTACLE (beta_cell's own stores);
//  * by the owner of x = y - c_l when x is OBSTACLE, with the post-alpha value
//    of (y, l) -- in place, nothing overwrites it during beta;
//  * by the owner of y when y - c_l lies outside the array (y is then an
//    OBSTACLE boundary cell, whose slots alpha leaves unchanged).
// This equals the in-place alpha + beta bit for bit provided every cell on
// the array boundary is OBSTACLE (no beta neighbour index wraps) and no
// rho / u outputs are stored; the host checks both before enabling the pass.
#pragma once

#include "lbm_cell.hpp"

namespace lbm {

// Column tile of TX x TY cells; the window adds a one-cell ring.
template <int TX, int TY>
struct PairShape {
    static constexpr int kTX = TX, kTY = TY;
    static constexpr int kWX = TX + 2;
    static constexpr int kWin = (TX + 2) * (TY + 2);
    static constexpr int kBlock = (kWin + 63) / 64 * 64;   // one window cell per thread
};
// the shapes the batch can launch (tuning.pairs = 1 + index)
using PairShape0 = PairShape<32, 8>;
using PairShape1 = PairShape<32, 4>;
using PairShape2 = PairShape<16, 8>;
using PairShape3 = PairShape<32, 6>;
constexpr int kPairShapes = 4;
constexpr int kPairZChunk = 64;                       // target beta planes per workgroup

struct PairDesc {
    const char* buf[2];   // the two population buffers
    const uint8_t* flags;
    int64_t stride;
    int32_t nx, ny, nz;
    int32_t tiles_x, tiles_y, chunks;
    uint32_t tile_begin;
    double inv_tau, lid, g0, g1, g2;
    uint32_t* status;
};

constexpr int kOutside = 0x100;                       // window cell outside the array

// Index of slot s inside its c_s.z group (+1: 5 slots, 0: 9, -1: 5).
__host__ __device__ constexpr int group_index(int s)
{
    constexpr int v[kQ] = {0, 1, 2, 3, 4, 5, 6, 7, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 8};
    return v[s];
}

// group_index is a bijection onto 0 .. size - 1 within each group
constexpr bool group_index_ok()
{
    for (int g = -1; g <= 1; g++) {
        const int size = g == 0 ? 9 : 5;
        int seen = 0, count = 0;
        for (int s = 0; s < kQ; s++)
            if (cz(s) == g) {
                if (group_index(s) < 0 || group_index(s) >= size || (seen >> group_index(s)) & 1)
                    return false;
                seen |= 1 << group_index(s);
                count++;
            }
        if (count != size)
            return false;
    }
    return true;
}
static_assert(group_index_ok(), "slot groups");

// LDS home of slot s of window plane p.
template <class S, int kWin>
struct PairLds {
    S* up;    // [5][kWin]       c.z = +1, plane p only
    S* mid;   // [2][9][kWin]    c.z =  0, planes p and p + 1
    S* down;  // [3][5][kWin]    c.z = -1, planes p .. p + 2
    __device__ __forceinline__ S* slot(int s, int p) const
    {
        if (cz(s) > 0)
            return up + group_index(s) * kWin;
        if (cz(s) == 0)
            return mid + (((p + 2) & 1) * 9 + group_index(s)) * kWin;
        return down + (((p + 3) % 3) * 5 + group_index(s)) * kWin;
    }
};

// Alpha of one window cell: inputs from registers, outputs to LDS.
template <class S, int kWin>
struct PairAlphaAcc {
    const S* in;  // the cell's 19 prefetched slots
    PairLds<S, kWin> lds;
    int p, c;
    __device__ __forceinline__ S load(int l, int32_t) const { return in[l]; }
    __device__ __forceinline__ void store(int s, int32_t, S v) const { lds.slot(s, p)[c] = v; }
};

// Beta of one tile cell: inputs from LDS (slot l of the window cell c + c_l),
// outputs to the destination buffer at id + c_l.
template <class S, bool NTS, int kWin, int kWX>
struct PairBetaAcc {
    PairLds<S, kWin> lds;
    int z, c;
    S* dst;
    int64_t stride;
    uint32_t id;
    __device__ __forceinline__ S load(int l, int32_t) const
    {
        return lds.slot(l, z + cz(l))[c + cy(l) * kWX + cx(l)];
    }
    __device__ __forceinline__ void store(int l, int32_t delta, S v) const
    {
        S* p = dst + (int64_t)l * stride + (int64_t)id + delta;
        if (NTS)
            __builtin_nontemporal_store(v, p);
        else
            *p = v;
    }
};

template <class C>
__device__ __forceinline__ StepArgs<C> pair_args(const PairDesc& d)
{
    StepArgs<C> a;
    a.dd = nullptr;
    a.flags = d.flags;
    a.rho = nullptr;
    a.u = nullptr;
    a.stride = d.stride;
    a.n = d.nx * d.ny * d.nz;
    a.nx = d.nx;
    a.nxy = d.nx * d.ny;
    a.inv_tau = (C)d.inv_tau;
    a.lid = (C)d.lid;
    a.g0 = (C)d.g0;
    a.g1 = (C)d.g1;
    a.g2 = (C)d.g2;
    a.store_rho = 0;
    a.store_u = 0;
    a.xpack_on = 0;
    a.xpack[0] = a.xpack[1] = nullptr;
    a.xunpack[0] = a.xunpack[1] = nullptr;
    a.status = d.status;
    return a;
}

template <class P, bool NTS, class G>
__global__ __launch_bounds__(G::kBlock) void k_pair(const PairDesc* __restrict__ mem,
                                                    const int32_t* __restrict__ tile_member,
                                                    uint32_t tiles, int parity, int xcd)
{
    using C = typename P::C;
    using S = typename P::S;
    constexpr int kPairTX = G::kTX, kPairTY = G::kTY, kWX = G::kWX, kWin = G::kWin;
    __shared__ S lds_up[5 * kWin];
    __shared__ S lds_mid[2 * 9 * kWin];
    __shared__ S lds_down[3 * 5 * kWin];
    __shared__ uint8_t lds_flag[3 * kWin];
    const PairLds<S, kWin> lds{lds_up, lds_mid, lds_down};

    const uint32_t t = xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    if (t >= tiles)
        return;
    const PairDesc& d = mem[tile_member[t]];
    uint32_t r = t - d.tile_begin;
    const int tx = (int)(r % (uint32_t)d.tiles_x);
    r /= (uint32_t)d.tiles_x;
    const int ty = (int)(r % (uint32_t)d.tiles_y);
    const int tc = (int)(r / (uint32_t)d.tiles_y);
    const int x0 = tx * kPairTX, y0 = ty * kPairTY;
    const int zb = (int)((int64_t)tc * d.nz / d.chunks);
    const int ze = (int)((int64_t)(tc + 1) * d.nz / d.chunks);
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const S* __restrict__ src = (const S*)d.buf[parity];
    S* __restrict__ dst = (S*)d.buf[parity ^ 1];
    const int64_t stride = d.stride;
    const StepArgs<C> a = pair_args<C>(d);

    // this thread's window cell (alpha) ...
    const int c = threadIdx.x;
    const int wy = c / kWX, wx = c - wy * kWX;
    const int ax = x0 - 1 + wx, ay = y0 - 1 + wy;
    const bool win_in = c < kWin && ax >= 0 && ax < nx && ay >= 0 && ay < ny;
    // ... and tile cell (beta)
    const int lx = c % kPairTX, ly = c / kPairTX;
    const int bx = x0 + lx, by = y0 + ly;
    const int cw = (ly + 1) * kWX + lx + 1;
    const bool tile_in = c < kPairTX * kPairTY && bx < nx && by < ny;

    // populations and flag of the window cell of plane p, into registers
    auto fetch = [&](int p, S (&s0)[kQ], int& flag) {
        flag = kOutside;
        if (win_in && p >= 0 && p < nz) {
            const int64_t id = ((int64_t)p * ny + ay) * nx + ax;
#pragma unroll
            for (int l = 0; l < kQ; l++)
                s0[l] = src[(int64_t)l * stride + id];
            flag = d.flags[id];
        }
    };
    // post-alpha slots and flag of the window cell of plane p into LDS
    auto alpha = [&](int p, const S (&s0)[kQ], int flag) {
        if (c < kWin)
            lds_flag[((p + 3) % 3) * kWin + c] = (uint8_t)(flag == kOutside ? 0 : flag);
        if (flag == kOutside)
            return;
        if (flag == kFluid || flag == kVelocityInjection) {
            const uint32_t id = (uint32_t)(((int64_t)p * ny + ay) * nx + ax);
            const PairAlphaAcc<S, kWin> acc{s0, lds, p, c};
            alpha_cell<P, kXfNone>(a, id, flag, acc);
        } else {
            // OBSTACLE / GHOST_LAYER: alpha leaves the slots as they are
#pragma unroll
            for (int s = 0; s < kQ; s++)
                lds.slot(s, p)[c] = s0[s];
        }
    };
    auto put = [&](int64_t i, S v) {
        if (NTS)
            __builtin_nontemporal_store(v, dst + i);
        else
            dst[i] = v;
    };
    auto beta = [&](int z) {
        if (!tile_in)
            return;
        const uint32_t id = (uint32_t)(((int64_t)z * ny + by) * nx + bx);
        const int flag = lds_flag[((z + 3) % 3) * kWin + cw];
        if (flag != kObstacle) {
            const PairBetaAcc<S, NTS, kWin, kWX> acc{lds, z, cw, dst, stride, id};
            beta_cell<P, kXfNone>(a, id, flag, acc);
            return;
        }
        // OBSTACLE: nothing overwrites the neighbours' post-alpha slots in place
#pragma unroll
        for (int l = 0; l < kQ; l++) {
            const int qx = bx + cx(l), qy = by + cy(l), qz = z + cz(l);
            if (qx < 0 || qx >= nx || qy < 0 || qy >= ny || qz < 0 || qz >= nz)
                continue;
            put((int64_t)l * stride + (int64_t)id + cx(l) + cy(l) * nx + (int64_t)cz(l) * nx * ny,
                lds.slot(l, qz)[cw + cy(l) * kWX + cx(l)]);
        }
        if (bx == 0 || by == 0 || z == 0 || bx == nx - 1 || by == ny - 1 || z == nz - 1) {
            // slots whose source cell lies outside the array keep their value
#pragma unroll
            for (int l = 0; l < kQ; l++) {
                const int sx = bx - cx(l), sy = by - cy(l), sz = z - cz(l);
                if (sx >= 0 && sx < nx && sy >= 0 && sy < ny && sz >= 0 && sz < nz)
                    continue;
                put((int64_t)l * stride + id, src[(int64_t)l * stride + id]);
            }
        }
    };

    // The next plane but one is fetched while alpha and beta of the current
    // planes run (one register set: a second costs more occupancy than it hides).
    S s0[kQ];
    int f0 = kOutside;
    fetch(zb - 1, s0, f0);
    alpha(zb - 1, s0, f0);
    fetch(zb, s0, f0);
    alpha(zb, s0, f0);
    fetch(zb + 1, s0, f0);
    for (int z = zb; z < ze; z++) {
        __syncthreads();          // beta of z - 1 is done with the slots alpha overwrites
        alpha(z + 1, s0, f0);
        if (z + 2 <= ze)
            fetch(z + 2, s0, f0); // in flight during beta below
        __syncthreads();
        beta(z);
    }
}

// 1 when some cell on the array boundary is not OBSTACLE.
__global__ void k_boundary_check(const uint8_t* __restrict__ flags, int nx, int ny, int nz,
                                 int* bad)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int x = (int)(i % nx);
        const int64_t t = i / nx;
        const int y = (int)(t % ny), z = (int)(t / ny);
        const bool edge = x == 0 || y == 0 || z == 0 || x == nx - 1 || y == ny - 1 || z == nz - 1;
        if (edge && flags[i] != kObstacle)
            *bad = 1;
    }
}

}  // namespace lbm
