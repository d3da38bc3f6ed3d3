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

#include "lbm_cell.hpp"

namespace lbm {

// ------------------------------------------------------------------ init ---

// lbm_init.cu:41-208 / CLbmInitCPU.cpp:37-171: flags from the six face
// conditions in x-fastest order (the first OBSTACLE wins, GHOST_LAYER can be
// overwritten), populations at the rest equilibrium computed with the
// reference's equilibrium arithmetic, u = 0, rho = 1.
template <class P>
__global__ __launch_bounds__(kBlock) void k_init(StepArgs<typename P::C> a, int ny,
                                                 int nz, int4 bc_lo, int2 bc_hi,
                                                 int write_flags)
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
    if (write_flags)
        const_cast<uint8_t*>(a.flags)[id] = (uint8_t)flag;
    if (id == 0)   // the rest state is in range for every storage format
        *a.status = 0u;

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

// Workgroup shapes: kShapeBase one cell per thread of a 256-thread workgroup;
// kShapeWide (LBM_LAUNCH_WIDE) 512-thread workgroups over 1024 consecutive
// cells, two cells per thread one block apart; kShapeBig (LBM_LAUNCH_BIG)
// 1024-thread workgroups, one cell per thread.
constexpr int kShapeBase = 0, kShapeWide = 1, kShapeBig = 2;
template <int SHAPE>
struct Shape {
    static constexpr uint32_t kThreads = SHAPE == kShapeWide  ? 2 * kBlock
                                         : SHAPE == kShapeBig ? 4 * kBlock
                                                              : kBlock;
    static constexpr uint32_t kCells = SHAPE == kShapeWide ? 2 : 1;   // per thread
    static constexpr uint32_t kSpan = kThreads * kCells;              // per workgroup
};

template <class P, bool FULL, bool OFF32, bool NT, int SHAPE, int XF>
__global__ __launch_bounds__(Shape<SHAPE>::kThreads) void k_alpha(StepArgs<typename P::C> a)
{
    using S = typename P::S;
    using G = Shape<SHAPE>;
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    for (uint32_t k = 0; k < G::kCells; k++) {
        const uint32_t gid = blk * G::kSpan + k * G::kThreads + threadIdx.x;
        if (gid >= a.cells)
            return;
        const uint32_t id = FULL ? gid : box_cell(a, gid);
        if (OFF32) {
            const Acc32<S, NT> acc{a.plane, id * (uint32_t)sizeof(S)};
            alpha_cell<P, XF>(a, id, a.flags[id], acc);
        } else {
            const Acc64<S, false> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
            alpha_cell<P, XF>(a, id, a.flags[id], acc);
        }
    }
}

template <class P, bool FULL, bool OFF32, bool NT, int SHAPE, int XF>
__global__ __launch_bounds__(Shape<SHAPE>::kThreads) void k_beta(StepArgs<typename P::C> a)
{
    using S = typename P::S;
    using G = Shape<SHAPE>;
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    // One decision per workgroup: neighbours of every cell of this block stay
    // inside [0, n) unless the block touches the first or last slab (box_cell
    // is increasing in gid, so the first and last cell bound the others).
    const uint32_t g_first = blk * G::kSpan;
    const uint32_t g_last = min(g_first + G::kSpan - 1, a.cells - 1);
    const uint32_t id_first = FULL ? g_first : box_cell(a, g_first);
    const uint32_t id_last = FULL ? g_last : box_cell(a, g_last);
    const int64_t reach = (int64_t)a.nxy + a.nx + 1;
    const bool wrap = (int64_t)id_first < reach || (int64_t)id_last + reach >= (int64_t)a.n;
    for (uint32_t k = 0; k < G::kCells; k++) {
        const uint32_t gid = g_first + k * G::kThreads + threadIdx.x;
        if (gid >= a.cells)
            return;
        const uint32_t id = FULL ? gid : box_cell(a, gid);
        if (wrap) {
            const Acc64<S, true> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
            beta_cell<P, XF>(a, id, a.flags[id], acc);
        } else if (OFF32) {
            const Acc32<S, NT> acc{a.plane, id * (uint32_t)sizeof(S)};
            beta_cell<P, XF>(a, id, a.flags[id], acc);
        } else {
            const Acc64<S, false> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
            beta_cell<P, XF>(a, id, a.flags[id], acc);
        }
    }
}

// Beta with its plane accesses staged through LDS (LBM_LAUNCH_LDS). Plane l
// of the cells [b, b + kBlock) lives at elements [b + s_l, b + s_l + kBlock)
// with s_l = c_l . (1, nx, nx * ny): a range that starts mid cache line for
// every moving direction. Each workgroup loads the 19 ranges from their
// 128-byte line starts (every wave instruction then covers whole lines),
// keeps them in LDS while every cell gathers, collides and scatters there
// (a cell reads and writes only its own entries, as in place), and stores the
// ranges back with the same aligned pattern, leaving out the elements of the
// lines' other cells. Blocks near the ends of the array (the linear wrap) and
// box steps use the per-cell path.
template <class S>
struct AccLds {
    S (*tile)[kBlock];
    uint32_t t;
    __device__ __forceinline__ S load(int l, int32_t) const { return tile[l][t]; }
    __device__ __forceinline__ void store(int l, int32_t, S v) const { tile[l][t] = v; }
};

template <class P, bool NT, int XF>
__global__ __launch_bounds__(kBlock) void k_beta_lds(StepArgs<typename P::C> a)
{
    using S = typename P::S;
    __shared__ S tile[kQ][kBlock];
    const uint32_t blk = a.xcd ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const uint32_t b = blk * kBlock, t = threadIdx.x;
    const uint32_t last = min(b + kBlock - 1, a.cells - 1);
    const int64_t reach = (int64_t)a.nxy + a.nx + 1;
    if ((int64_t)b < reach || (int64_t)last + reach >= (int64_t)a.n || last != b + kBlock - 1) {
        // uniform per workgroup: no barrier on this path
        const uint32_t id = b + t;
        if (id < a.cells) {
            const Acc64<S, true> acc{(S*)a.dd, a.stride, (int32_t)id, a.n};
            beta_cell<P, XF>(a, id, a.flags[id], acc);
        }
        return;
    }
    // the 19 ranges, each from the start of its first line: element j of the
    // aligned run is tile position j - mis, valid for mis <= j < kBlock + mis
    S head[kQ], tail[kQ];
    uint32_t mis[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const char* base = a.plane[l] + (size_t)b * sizeof(S);
        mis[l] = (uint32_t)(((uintptr_t)base & 127) / sizeof(S));
        const LBM_GLOBAL S* run = (const LBM_GLOBAL S*)(base - mis[l] * sizeof(S));
        head[l] = NT ? __builtin_nontemporal_load(run + t) : run[t];
        if (t < mis[l])
            tail[l] = NT ? __builtin_nontemporal_load(run + kBlock + t) : run[kBlock + t];
    }
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        if (t >= mis[l])
            tile[l][t - mis[l]] = head[l];
        if (t < mis[l])
            tile[l][kBlock - mis[l] + t] = tail[l];
    }
    __syncthreads();
    const uint32_t id = b + t;
    beta_cell<P, XF>(a, id, a.flags[id], AccLds<S>{tile, t});
    __syncthreads();
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const char* base = a.plane[l] + (size_t)b * sizeof(S);
        LBM_GLOBAL S* run = (LBM_GLOBAL S*)(const_cast<char*>(base) - mis[l] * sizeof(S));
        if (t >= mis[l]) {
            if (NT)
                __builtin_nontemporal_store(tile[l][t - mis[l]], run + t);
            else
                run[t] = tile[l][t - mis[l]];
        }
        if (t < mis[l]) {
            if (NT)
                __builtin_nontemporal_store(tile[l][kBlock - mis[l] + t], run + kBlock + t);
            else
                run[kBlock + t] = tile[l][kBlock - mis[l] + t];
        }
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
    a.xunpack[0] = (const C*)s->xunpack[0];
    a.xunpack[1] = (const C*)s->xunpack[1];
    a.xpack_on = (s->xpack[0] || s->xpack[1] || s->xunpack[0] || s->xunpack[1]) ? 1 : 0;
    a.div_nx = FastDiv((uint32_t)s->size_h[0]);
    a.ny = s->size_h[1];
    a.fcells = s->size_h[1] * s->size_h[2];
    a.status = s->status;
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
static hipError_t init_t(const lbm_solver* s, hipStream_t st, int write_flags)
{
    Box all{0, 0, 0, s->size_h[0], s->size_h[1], s->size_h[2]};
    auto a = make_args<P>(s, all);
    const uint32_t blocks = (uint32_t)((s->n + kBlock - 1) / kBlock);
    int4 lo = make_int4(s->p.bc[0], s->p.bc[1], s->p.bc[2], s->p.bc[3]);
    int2 hi = make_int2(s->p.bc[4], s->p.bc[5]);
    hipLaunchKernelGGL(k_init<P>, dim3(blocks), dim3(kBlock), 0, st, a, s->size_h[1],
                       s->size_h[2], lo, hi, write_flags);
    return hipGetLastError();
}


// LBM_LAUNCH_WAVES4 / WAVES3 / WAVES2: reserving this much (unused) LDS per
// 256-thread workgroup leaves room for 4 (3, 2) workgroups in a CU's 160 KB,
// i.e. 4 (3, 2) waves per SIMD. Fewer resident waves means fewer concurrent
// 19-plane streams, which the in-place FP32 steps turn into higher DRAM
// efficiency.
constexpr size_t kWaves4Lds = 33 * 1024;
constexpr size_t kWaves3Lds = 41 * 1024;
constexpr size_t kWaves2Lds = 54 * 1024;

// Reserved LDS for a waves-per-SIMD cap with WIDE (512-thread) workgroups:
// 4 waves per SIMD leave room for 2 workgroups per CU, 3 or 2 for one. A BIG
// (1024-thread) workgroup alone fills 4 waves per SIMD: every cap holds one.
constexpr size_t kWideWaves4Lds = 54 * 1024;
constexpr size_t kWideWaves2Lds = 81 * 1024;

template <class P, template <class, bool, bool, bool, int, int> class K, bool FULL, int SHAPE,
          int XF>
static void launch_shape(const lbm_solver* s, const StepArgs<typename P::C>& a, int opts,
                         hipStream_t st)
{
    using G = Shape<SHAPE>;
    const uint32_t blocks = (uint32_t)(((uint64_t)a.cells + G::kSpan - 1) / G::kSpan);
    const bool cap2 = opts & (LBM_LAUNCH_WAVES2 | LBM_LAUNCH_WAVES3);
    const bool any_cap = opts & (LBM_LAUNCH_WAVES2 | LBM_LAUNCH_WAVES3 | LBM_LAUNCH_WAVES4);
    const size_t lds = SHAPE == kShapeBig  ? (any_cap ? kWideWaves2Lds : 0)
                       : SHAPE == kShapeWide ? (cap2                           ? kWideWaves2Lds
                                                : (opts & LBM_LAUNCH_WAVES4) ? kWideWaves4Lds
                                                                             : 0)
                       : (opts & LBM_LAUNCH_WAVES2) ? kWaves2Lds
                       : (opts & LBM_LAUNCH_WAVES3) ? kWaves3Lds
                       : (opts & LBM_LAUNCH_WAVES4) ? kWaves4Lds : 0;
    const dim3 grid(blocks), block(G::kThreads);
    if (!plane_off32(s) || (opts & LBM_LAUNCH_ADDR64))
        hipLaunchKernelGGL((K<P, FULL, false, false, SHAPE, XF>::kernel), grid, block, lds, st, a);
    else if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((K<P, FULL, true, true, SHAPE, XF>::kernel), grid, block, lds, st, a);
    else
        hipLaunchKernelGGL((K<P, FULL, true, false, SHAPE, XF>::kernel), grid, block, lds, st, a);
}

template <class P, template <class, bool, bool, bool, int, int> class K, bool FULL, int XF>
static void launch_xf(const lbm_solver* s, const StepArgs<typename P::C>& a, int opts,
                      hipStream_t st)
{
    if (opts & LBM_LAUNCH_BIG)
        launch_shape<P, K, FULL, kShapeBig, XF>(s, a, opts, st);
    else if (opts & LBM_LAUNCH_WIDE)
        launch_shape<P, K, FULL, kShapeWide, XF>(s, a, opts, st);
    else
        launch_shape<P, K, FULL, kShapeBase, XF>(s, a, opts, st);
}

// The face work a solver's steps do: the fused unpack (with or without a
// pack), the fused pack alone, or neither.
static int face_work(const lbm_solver* s)
{
    return (s->xunpack[0] || s->xunpack[1]) ? kXfUnpack
           : (s->xpack[0] || s->xpack[1])   ? kXfPack
                                            : kXfNone;
}

template <class P, template <class, bool, bool, bool, int, int> class K, bool FULL>
static void launch_variant(const lbm_solver* s, const StepArgs<typename P::C>& a, int opts,
                           hipStream_t st)
{
    switch (face_work(s)) {
    case kXfUnpack: launch_xf<P, K, FULL, kXfUnpack>(s, a, opts, st); break;
    case kXfPack: launch_xf<P, K, FULL, kXfPack>(s, a, opts, st); break;
    default: launch_xf<P, K, FULL, kXfNone>(s, a, opts, st); break;
    }
}

template <class P, bool FULL, bool OFF32, bool NT, int SHAPE, int XF>
struct AlphaK {
    static constexpr auto kernel = k_alpha<P, FULL, OFF32, NT, SHAPE, XF>;
};
template <class P, bool FULL, bool OFF32, bool NT, int SHAPE, int XF>
struct BetaK {
    static constexpr auto kernel = k_beta<P, FULL, OFF32, NT, SHAPE, XF>;
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

// LBM_LAUNCH_LDS beta: the static tile counts towards the LDS reservation of
// the waves-per-SIMD caps.
template <class P, int XF>
static void launch_beta_lds_xf(const StepArgs<typename P::C>& a, int opts, hipStream_t st)
{
    constexpr size_t kTile = (size_t)kQ * kBlock * sizeof(typename P::S);
    const size_t want = (opts & LBM_LAUNCH_WAVES2)   ? kWaves2Lds
                        : (opts & LBM_LAUNCH_WAVES3) ? kWaves3Lds
                        : (opts & LBM_LAUNCH_WAVES4) ? kWaves4Lds : 0;
    const size_t lds = want > kTile ? want - kTile : 0;
    const uint32_t blocks = (a.cells + kBlock - 1) / kBlock;
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_beta_lds<P, true, XF>), dim3(blocks), dim3(kBlock), lds, st, a);
    else
        hipLaunchKernelGGL((k_beta_lds<P, false, XF>), dim3(blocks), dim3(kBlock), lds, st, a);
}

template <class P>
static void launch_beta_lds(const lbm_solver* s, const StepArgs<typename P::C>& a, int opts,
                            hipStream_t st)
{
    switch (face_work(s)) {
    case kXfUnpack: launch_beta_lds_xf<P, kXfUnpack>(a, opts, st); break;
    case kXfPack: launch_beta_lds_xf<P, kXfPack>(a, opts, st); break;
    default: launch_beta_lds_xf<P, kXfNone>(a, opts, st); break;
    }
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
    const int opts = s->launch[1];
    if ((opts & LBM_LAUNCH_LDS) && is_full(s, b) && plane_off32(s) && !(opts & LBM_LAUNCH_ADDR64)) {
        launch_beta_lds<P>(s, a, opts, st);
        return hipGetLastError();
    }
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
    case LBM_F16C: return fn<StoreF16C>(__VA_ARGS__);          \
    default: return hipErrorInvalidValue;                      \
    }

hipError_t launch_init(const lbm_solver* s, hipStream_t st, int write_flags)
{
    LBM_DISPATCH(init_t, s, st, write_flags)
}
hipError_t launch_alpha(const lbm_solver* s, const Box& b, hipStream_t st)
{
    LBM_DISPATCH(alpha_t, s, b, st)
}
hipError_t launch_beta(const lbm_solver* s, const Box& b, hipStream_t st)
{
    LBM_DISPATCH(beta_t, s, b, st)
}

}  // namespace lbm
