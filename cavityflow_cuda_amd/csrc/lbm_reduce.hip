// FP64 reductions over the population field: total mass / momentum of a box
// and the steady-state residual ||u - u_prev||_2 / ||u||_2. Two passes with a
// fixed block count and a fixed final order, so results are deterministic.
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"
#include "lbm_storage.hpp"

#include <mutex>
#include <unordered_map>

namespace lbm {

constexpr int kRedBlock = 256;
constexpr int kRedVals = 5;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

// Block sum of kRedVals doubles into partials[block * kRedVals + k].
__device__ __forceinline__ void block_store(double (&acc)[kRedVals], double* partials,
                                            uint32_t block)
{
    __shared__ double lds[kRedBlock / 64][kRedVals];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < kRedVals; k++) {
        const double v = wave_sum(acc[k]);
        if (lane == 0)
            lds[wave][k] = v;
    }
    __syncthreads();
    if (threadIdx.x < kRedVals) {
        double t = 0.0;
        for (int w = 0; w < kRedBlock / 64; w++)
            t += lds[w][threadIdx.x];
        partials[block * kRedVals + threadIdx.x] = t;
    }
}

struct RedArgs {
    int64_t stride;
    int32_t nx, nxy;
    int32_t ox, oy, oz;
    uint32_t cells;
    FastDiv div_bx, div_by;
    const uint32_t* status;   // the solver's overflow word
};

// A state with a population store out of its storage range (F16C) has no
// meaningful sums: they become NaN, as a non-finite state's are.
__device__ __forceinline__ bool poisoned(const uint32_t* status)
{
    return status && *status != 0u;
}

// Block `block` of `nblocks` strides over the cells of the box; its sums go to
// partials[block].
template <class P>
__device__ __forceinline__ void reduce_box_block(const typename P::S* __restrict__ dd,
                                                 const RedArgs& a, uint32_t block,
                                                 uint32_t nblocks, double* partials)
{
    double acc[kRedVals] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t gid = block * kRedBlock + threadIdx.x; gid < a.cells;
         gid += nblocks * kRedBlock) {
        const uint32_t t = a.div_bx.div(gid);
        const uint32_t x = gid - t * a.div_bx.d;
        const uint32_t z = a.div_by.div(t);
        const uint32_t y = t - z * a.div_by.d;
        const uint32_t id = (uint32_t)((a.oz + (int32_t)z) * a.nxy + (a.oy + (int32_t)y) * a.nx +
                                       a.ox + (int32_t)x);
#pragma unroll
        for (int l = 0; l < kQ; l++) {
            const double f = (double)P::load(dd[(int64_t)l * a.stride + id], l);
            acc[0] += f;
            acc[1] += cx(l) * f;
            acc[2] += cy(l) * f;
            acc[3] += cz(l) * f;
        }
        acc[4] += 1.0;
    }
    if (poisoned(a.status))
        for (int k = 0; k < 4; k++)
            acc[k] = __builtin_nan("");
    block_store(acc, partials, block);
}

template <class P>
__global__ __launch_bounds__(kRedBlock) void k_reduce_box(const typename P::S* __restrict__ dd,
                                                          RedArgs a, double* partials)
{
    reduce_box_block<P>(dd, a, blockIdx.x, gridDim.x, partials);
}

// The box sums of several solvers in one launch: grid.y = member, and member
// m's blocks do exactly what k_reduce_box's do for that box alone, into their
// own slice of partials, so each member's sums carry lbm_reduce_box's bits.
struct ReduceBoxMember {
    const void* dd;
    RedArgs a;
};
struct ReduceBoxMembers {
    ReduceBoxMember m[LBM_NOALLOC_MAX_BOXES];
};

template <class P>
__global__ __launch_bounds__(kRedBlock) void k_reduce_boxes(ReduceBoxMembers mem,
                                                            double* partials)
{
    const ReduceBoxMember& m = mem.m[blockIdx.y];
    reduce_box_block<P>((const typename P::S*)m.dd, m.a, blockIdx.x, gridDim.x,
                        partials + (size_t)blockIdx.y * gridDim.x * kRedVals);
}

// Residual over FLUID cells. LAYOUT 0: slot l holds direction l (after beta);
// LAYOUT 1: slot l holds direction opp(l) (after alpha). A template argument:
// with a run-time layout the slot permutation became register-indexed moves
// that serialise the 19 loads. Block `block` of `nblocks` strides over the
// cells; its sums go to partials[block].
template <class P, int LAYOUT>
__device__ __forceinline__ void residual_block(const typename P::S* __restrict__ dd,
                                               const uint8_t* __restrict__ flags,
                                               typename P::C* __restrict__ prev,
                                               int64_t stride, uint32_t n,
                                               uint32_t block, uint32_t nblocks,
                                               double* partials, const uint32_t* status)
{
    using C = typename P::C;
    double acc[kRedVals] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (uint32_t id = block * kRedBlock + threadIdx.x; id < n; id += nblocks * kRedBlock) {
        // every load of the cell is issued before the flag decides (one memory
        // round trip per cell instead of two); cells other than FLUID add
        // +0.0, which leaves the non-negative sums' bits as skipping them
        // would, and store their snapshot back unchanged (no branch for the
        // compiler to sink the snapshot loads into)
        const bool fluid = flags[id] == kFluid;
        C d[kQ];
#pragma unroll
        for (int l = 0; l < kQ; l++) {
            d[LAYOUT ? opp(l) : l] = P::load(dd[(int64_t)l * stride + id], l);
        }
        const C p0 = prev[id], p1 = prev[stride + id], p2 = prev[2 * stride + id];
        // empty statements that read every loaded value: the compiler cannot
        // sink the loads below the flag test into the FLUID path
#pragma unroll
        for (int l = 0; l < kQ; l++)
            asm volatile("" ::"v"(d[l]));
        asm volatile("" ::"v"(p0), "v"(p1), "v"(p2));
        C rho, ux, uy, uz;
        moments(d, rho, ux, uy, uz);
        const double dx = (double)ux - (double)p0;
        const double dy = (double)uy - (double)p1;
        const double dz = (double)uz - (double)p2;
        acc[0] += fluid ? dx * dx + dy * dy + dz * dz : 0.0;
        acc[1] += fluid ? (double)ux * ux + (double)uy * uy + (double)uz * uz : 0.0;
        acc[2] += fluid ? 1.0 : 0.0;
        prev[id] = fluid ? ux : p0;
        prev[stride + id] = fluid ? uy : p1;
        prev[2 * stride + id] = fluid ? uz : p2;
    }
    if (poisoned(status))
        acc[0] = acc[1] = __builtin_nan("");
    block_store(acc, partials, block);
}

template <class P, int LAYOUT>
__global__ __launch_bounds__(kRedBlock) void k_residual(const typename P::S* __restrict__ dd,
                                                        const uint8_t* __restrict__ flags,
                                                        typename P::C* __restrict__ prev,
                                                        int64_t stride, uint32_t n,
                                                        double* partials,
                                                        const uint32_t* status)
{
    residual_block<P, LAYOUT>(dd, flags, prev, stride, n, blockIdx.x, gridDim.x, partials,
                              status);
}

// The residual of many solvers in one launch: grid.y = member, and member m's
// blocks do exactly what k_residual's do for that solver alone, into their own
// slice of partials, so each member's sums carry the same bits as
// lbm_residual's.
template <class P, int LAYOUT>
__global__ __launch_bounds__(kRedBlock) void k_residual_members(const ResidualMember* mem,
                                                                double* partials)
{
    const ResidualMember& m = mem[blockIdx.y];
    residual_block<P, LAYOUT>((const typename P::S*)m.dd, m.flags, (typename P::C*)m.prev,
                              m.stride, m.n, blockIdx.x, gridDim.x,
                              partials + (size_t)blockIdx.y * gridDim.x * kRedVals, m.status);
}

// Density and momentum of every cell of a box, computed from the populations
// with the reference accumulation order; dst is dense [rho, jx, jy, jz][box].
template <class P, int LAYOUT>
__global__ __launch_bounds__(kRedBlock) void k_moments(const typename P::S* __restrict__ dd,
                                                       RedArgs a,
                                                       typename P::C* __restrict__ dst)
{
    using C = typename P::C;
    const uint32_t gid = blockIdx.x * kRedBlock + threadIdx.x;
    if (gid >= a.cells)
        return;
    const uint32_t t = a.div_bx.div(gid);
    const uint32_t x = gid - t * a.div_bx.d;
    const uint32_t z = a.div_by.div(t);
    const uint32_t y = t - z * a.div_by.d;
    const uint32_t id = (uint32_t)((a.oz + (int32_t)z) * a.nxy + (a.oy + (int32_t)y) * a.nx +
                                   a.ox + (int32_t)x);
    C d[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        d[LAYOUT ? opp(l) : l] = P::load(dd[(int64_t)l * a.stride + id], l);
    C rho, ux, uy, uz;
    moments(d, rho, ux, uy, uz);
    if (poisoned(a.status))
        rho = ux = uy = uz = (C)__builtin_nan("");
    dst[gid] = rho;
    dst[(int64_t)a.cells + gid] = ux;
    dst[2 * (int64_t)a.cells + gid] = uy;
    dst[3 * (int64_t)a.cells + gid] = uz;
}

// Sum of the per-block partials in a fixed order (lane i adds blocks i, i + 64,
// ... in turn, then the butterfly over the wave): the same bits on every call,
// in a few microseconds instead of one lane walking all blocks.
// grid.x > 1: one slice of nblocks partials and one output per workgroup.
__global__ __launch_bounds__(64) void k_finalize(const double* partials, int nblocks,
                                                 double* out)
{
    partials += (size_t)blockIdx.x * nblocks * kRedVals;
    out += (size_t)blockIdx.x * kRedVals;
#pragma unroll
    for (int k = 0; k < kRedVals; k++) {
        double t = 0.0;
        for (int b = threadIdx.x; b < nblocks; b += 64)
            t += partials[b * kRedVals + k];
        t = wave_sum(t);
        if (threadIdx.x == 0)
            out[k] = t;
    }
}

static RedArgs red_args(const lbm_solver* s, const Box& b)
{
    RedArgs a;
    a.stride = s->stride;
    a.nx = s->size_h[0];
    a.nxy = s->size_h[0] * s->size_h[1];
    a.ox = b.ox;
    a.oy = b.oy;
    a.oz = b.oz;
    a.cells = b.cells();
    a.div_bx = FastDiv((uint32_t)b.bx);
    a.div_by = FastDiv((uint32_t)b.by);
    a.status = s->status;
    return a;
}

template <class P>
static hipError_t reduce_box_t(const lbm_solver* s, const Box& b, double* partials,
                               double* out, hipStream_t st)
{
    const RedArgs a = red_args(s, b);
    hipLaunchKernelGGL(k_reduce_box<P>, dim3(kReduceBlocks), dim3(kRedBlock), 0, st,
                       (const typename P::S*)s->dd, a, partials);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, st, partials, kReduceBlocks, out);
    return hipGetLastError();
}

template <class P>
static hipError_t residual_t(const lbm_solver* s, int layout, double* partials, double* out,
                             hipStream_t st)
{
    if (layout)
        hipLaunchKernelGGL((k_residual<P, 1>), dim3(kReduceBlocks), dim3(kRedBlock), 0, st,
                           (const typename P::S*)s->dd, s->flags,
                           (typename P::C*)s->residual_prev, s->stride, (uint32_t)s->n,
                           partials, s->status);
    else
        hipLaunchKernelGGL((k_residual<P, 0>), dim3(kReduceBlocks), dim3(kRedBlock), 0, st,
                           (const typename P::S*)s->dd, s->flags,
                           (typename P::C*)s->residual_prev, s->stride, (uint32_t)s->n,
                           partials, s->status);
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, st, partials, kReduceBlocks, out);
    return hipGetLastError();
}

template <class P>
static hipError_t moments_t(const lbm_solver* s, const Box& b, int layout, void* dst,
                            hipStream_t st)
{
    const RedArgs a = red_args(s, b);
    const dim3 grid((a.cells + kRedBlock - 1) / kRedBlock);
    if (layout)
        hipLaunchKernelGGL((k_moments<P, 1>), grid, dim3(kRedBlock), 0, st,
                           (const typename P::S*)s->dd, a, (typename P::C*)dst);
    else
        hipLaunchKernelGGL((k_moments<P, 0>), grid, dim3(kRedBlock), 0, st,
                           (const typename P::S*)s->dd, a, (typename P::C*)dst);
    return hipGetLastError();
}

hipError_t launch_moments(const lbm_solver* s, const Box& b, int layout, void* dst,
                          hipStream_t st)
{
    switch (s->p.precision) {
    case LBM_F32: return moments_t<StoreF32>(s, b, layout, dst, st);
    case LBM_F64: return moments_t<StoreF64>(s, b, layout, dst, st);
    case LBM_F16S: return moments_t<StoreF16S>(s, b, layout, dst, st);
    case LBM_BF16S: return moments_t<StoreBF16S>(s, b, layout, dst, st);
    case LBM_F16C: return moments_t<StoreF16C>(s, b, layout, dst, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_reduce_box(const lbm_solver* s, const Box& b, double* partials,
                             double* out_dev, hipStream_t st)
{
    switch (s->p.precision) {
    case LBM_F32: return reduce_box_t<StoreF32>(s, b, partials, out_dev, st);
    case LBM_F64: return reduce_box_t<StoreF64>(s, b, partials, out_dev, st);
    case LBM_F16S: return reduce_box_t<StoreF16S>(s, b, partials, out_dev, st);
    case LBM_BF16S: return reduce_box_t<StoreBF16S>(s, b, partials, out_dev, st);
    case LBM_F16C: return reduce_box_t<StoreF16C>(s, b, partials, out_dev, st);
    default: return hipErrorInvalidValue;
    }
}

template <class P>
static hipError_t residual_members_t(const ResidualMember* mem, int count, int layout,
                                     double* partials, double* out, hipStream_t st)
{
    if (layout)
        hipLaunchKernelGGL((k_residual_members<P, 1>), dim3(kReduceBlocks, count),
                           dim3(kRedBlock), 0, st, mem, partials);
    else
        hipLaunchKernelGGL((k_residual_members<P, 0>), dim3(kReduceBlocks, count),
                           dim3(kRedBlock), 0, st, mem, partials);
    hipLaunchKernelGGL(k_finalize, dim3(count), dim3(64), 0, st, partials, kReduceBlocks, out);
    return hipGetLastError();
}

hipError_t launch_residual_members(int precision, const ResidualMember* mem, int count,
                                   int layout, double* partials, double* out_dev,
                                   hipStream_t st)
{
    switch (precision) {
    case LBM_F32: return residual_members_t<StoreF32>(mem, count, layout, partials, out_dev, st);
    case LBM_F64: return residual_members_t<StoreF64>(mem, count, layout, partials, out_dev, st);
    case LBM_F16S:
        return residual_members_t<StoreF16S>(mem, count, layout, partials, out_dev, st);
    case LBM_BF16S:
        return residual_members_t<StoreBF16S>(mem, count, layout, partials, out_dev, st);
    case LBM_F16C:
        return residual_members_t<StoreF16C>(mem, count, layout, partials, out_dev, st);
    default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------ allocation-free sums ---

// Block partials of lbm_reduce_boxes_noalloc: LBM_NOALLOC_MAX_BOXES slices of
// kReduceBlocks * kRedVals doubles (640 KiB). Every GPU the library runs on
// holds its own copy until the process ends. In host code `scratch` names a
// host-side stand-in, not a device address, so it is never passed to a kernel
// or a copy: each device's address comes from hipGetSymbolAddress, looked up
// once per device with that device current (noalloc_scratch). Every block
// overwrites its whole slice, so nothing has to start at zero; two calls that
// overlap on one GPU share the array and race.
static __device__ __align__(256) double scratch[LBM_NOALLOC_MAX_BOXES * kReduceBlocks * kRedVals];

static std::mutex g_scratch_mutex;
static std::unordered_map<int, double*> g_scratch_ptr;

hipError_t noalloc_scratch(int device, double** ptr)
{
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *ptr = it->second;
        return hipSuccess;
    }
    // the caller has made `device` current
    double* p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
        return e;
    g_scratch_ptr.emplace(device, p);
    *ptr = p;
    return hipSuccess;
}

template <class P>
static hipError_t reduce_boxes_t(const lbm_solver* const* s, const Box* b, int count,
                                 double* partials, double* out, hipStream_t st)
{
    ReduceBoxMembers mem;
    for (int m = 0; m < count; m++) {
        mem.m[m].dd = s[m]->dd;
        mem.m[m].a = red_args(s[m], b[m]);
    }
    hipLaunchKernelGGL(k_reduce_boxes<P>, dim3(kReduceBlocks, count), dim3(kRedBlock), 0, st,
                       mem, partials);
    hipLaunchKernelGGL(k_finalize, dim3(count), dim3(64), 0, st, partials, kReduceBlocks, out);
    return hipGetLastError();
}

hipError_t launch_reduce_boxes(const lbm_solver* const* s, const Box* b, int count,
                               double* partials, double* out_dev, hipStream_t st)
{
    if (count < 1 || count > LBM_NOALLOC_MAX_BOXES)
        return hipErrorInvalidValue;
    switch (s[0]->p.precision) {
    case LBM_F32: return reduce_boxes_t<StoreF32>(s, b, count, partials, out_dev, st);
    case LBM_F64: return reduce_boxes_t<StoreF64>(s, b, count, partials, out_dev, st);
    case LBM_F16S: return reduce_boxes_t<StoreF16S>(s, b, count, partials, out_dev, st);
    case LBM_BF16S: return reduce_boxes_t<StoreBF16S>(s, b, count, partials, out_dev, st);
    case LBM_F16C: return reduce_boxes_t<StoreF16C>(s, b, count, partials, out_dev, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_residual(const lbm_solver* s, int layout, double* partials, double* out_dev,
                           hipStream_t st)
{
    switch (s->p.precision) {
    case LBM_F32: return residual_t<StoreF32>(s, layout, partials, out_dev, st);
    case LBM_F64: return residual_t<StoreF64>(s, layout, partials, out_dev, st);
    case LBM_F16S: return residual_t<StoreF16S>(s, layout, partials, out_dev, st);
    case LBM_BF16S: return residual_t<StoreBF16S>(s, layout, partials, out_dev, st);
    case LBM_F16C: return residual_t<StoreF16C>(s, layout, partials, out_dev, st);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace lbm
