// Internal declarations of the HIP solver (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <limits>
#include <string>

#include "lbm_abi.h"

namespace lbm {

// Unsigned 32-bit division by a runtime constant: q = (umulhi(n, m) + n) >> s,
// exact for every 32-bit n (round-up multiplier, Granlund & Montgomery 1994).
struct FastDiv {
    uint32_t d = 1, m = 0, s = 0;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div)
    {
        uint32_t l = 0;
        while (l < 32 && (1ull << l) < div)
            l++;
        s = l;
        m = (uint32_t)((((1ull << l) - div) << 32) / div + 1);
        if (div == 1)
            m = 0, s = 0;
    }
    __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        return (uint32_t)(((uint64_t)__umulhi(n, m) + n) >> s);
    }
};

// A box of cells in halo coordinates.
struct Box {
    int ox, oy, oz;
    int bx, by, bz;
    uint32_t cells() const { return (uint32_t)bx * (uint32_t)by * (uint32_t)bz; }
};

// Everything a step kernel needs, passed by value.
template <class C>
struct StepArgs {
    void* dd;
    const uint8_t* flags;
    C* rho;
    C* u;
    int64_t stride;   // elements between planes
    int32_t n;        // cells with halo
    int32_t nx, nxy;  // row and slab strides
    int32_t ox, oy, oz;
    uint32_t cells;   // cells of the box
    FastDiv div_bx, div_by;
    C inv_tau, lid, g0, g1, g2;
    int store_rho, store_u;
    int xcd;          // renumber workgroups per XCD
    // byte address of plane l, shifted by the neighbour offset of direction l
    // for beta (0 for alpha): uniform per-plane bases for 32-bit lane offsets
    const char* plane[19];
    // fused x-face pack: dense face buffers [slot][z * ny + y] (LEFT, RIGHT)
    C* xpack[2];
    // fused x-face unpack: the received face buffers, same layout (LEFT, RIGHT)
    const C* xunpack[2];
    int xpack_on;     // any of xpack / xunpack set
    FastDiv div_nx;
    int32_t ny;
    int32_t fcells;   // ny * nz
    // set to 1 by a store the storage format cannot represent (F16C)
    uint32_t* status;
};

}  // namespace lbm

struct lbm_solver {
    lbm_params p;
    int size_h[3];
    int64_t n;         // cells with halo
    int64_t stride;    // plane stride in elements
    int elem_bytes;    // bytes per stored population
    int out_bytes;     // bytes per rho/u element (4 or 8)
    void* dd;
    uint8_t* flags;
    void* rho;
    void* u;
    void* staging;     // device staging for host copies
    size_t staging_bytes;
    double* partials;  // reduction partials (kReduceBlocks * 5) followed by 5 results
    void* residual_prev;  // momentum saved by lbm_residual, 3 planes of the compute type
    int residual_valid;
    int launch[2];     // LBM_LAUNCH_* options of the alpha and beta kernels
    int launch_set[2]; // set by the caller: kept when the step pack changes
    void* xpack[2];    // fused x-face pack buffers (LEFT, RIGHT), NULL when off
    const void* xunpack[2];  // fused x-face unpack buffers (LEFT, RIGHT), NULL when off
    void* dd_alt;      // second population buffer of the out-of-place pair pass
    uint32_t* status;  // device word: nonzero after a store out of the storage range
};

namespace lbm {

void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

hipError_t launch_init(const lbm_solver* s, hipStream_t st, int write_flags);
hipError_t launch_alpha(const lbm_solver* s, const Box& b, hipStream_t st);
hipError_t launch_beta(const lbm_solver* s, const Box& b, hipStream_t st);

// field: 0 populations (the lattice vectors of `mask`; buffers keep all 19
// planes), 1 velocities (3), 2 densities (1)
hipError_t launch_box_get(const lbm_solver* s, int field, const Box& b, void* dst,
                          uint32_t mask, hipStream_t st);
// the population copies of two boxes (the faces of one axis) in one launch,
// each with its own buffer and plane mask
hipError_t launch_dd_get_pair(const lbm_solver* s, const Box* b, void* const* dst,
                              const uint32_t* mask, hipStream_t st);
// faces of one solver copied into another's with no buffer
// (lbm_exchange_face[s]): the boxes of a transfer have the same extents, the
// solvers one device and precision, and no element is written by one
// transfer of the list and read or written by another
struct XferDesc {
    const lbm_solver* from;
    Box sb;
    lbm_solver* to;
    Box db;
    uint32_t mask;
};
constexpr int kXferBatch = 8;   // transfers per launch
hipError_t launch_dd_xfer(const XferDesc* list, int n, hipStream_t st);
hipError_t launch_dd_set_pair(const lbm_solver* s, const Box* b, const void* const* src,
                              const uint32_t* mask, hipStream_t st);
// both x faces in one launch (nt: streaming stores)
hipError_t launch_set_dd_xpair(const lbm_solver* s, const void* lo, const void* hi, int x_lo,
                               int x_hi, uint32_t mask_lo, uint32_t mask_hi, bool nt,
                               hipStream_t st);
hipError_t launch_box_set(const lbm_solver* s, int field, const Box& b, const void* src,
                          uint32_t mask, hipStream_t st);
hipError_t launch_flags_get(const lbm_solver* s, const Box& b, int32_t* dst, hipStream_t st);
hipError_t launch_flags_set(const lbm_solver* s, const Box& b, const int32_t* src,
                            hipStream_t st);

hipError_t launch_reduce_box(const lbm_solver* s, const Box& b, double* partials,
                             double* out_dev, hipStream_t st);
hipError_t launch_moments(const lbm_solver* s, const Box& b, int layout, void* dst,
                          hipStream_t st);
hipError_t launch_residual(const lbm_solver* s, int layout, double* partials, double* out_dev,
                           hipStream_t st);
// One solver's residual inputs, for the residual of many solvers in one launch
// (partials: count * kReduceBlocks * 5 doubles; out: count * 5).
struct ResidualMember {
    const void* dd;
    const uint8_t* flags;
    void* prev;
    int64_t stride;
    uint32_t n;
    const uint32_t* status;
};
hipError_t launch_residual_members(int precision, const ResidualMember* mem, int count,
                                   int layout, double* partials, double* out_dev,
                                   hipStream_t st);

constexpr int kReduceBlocks = 1024;
// Per-solver reduction buffer: kReduceBlocks * 5 partial sums, 5 results,
// then the solver's overflow word (lbm_solver::status).
constexpr size_t kPartialsBytes = (size_t)(kReduceBlocks + 1) * 5 * sizeof(double) + 64;

// lbm_reduce_boxes_noalloc: the device address of the static partials array
// on `device` (which must be current), looked up on the first call per device.
hipError_t noalloc_scratch(int device, double** ptr);
// k_reduce_box for count (<= LBM_NOALLOC_MAX_BOXES) boxes of solvers of one
// precision; partials: count * kReduceBlocks * 5 doubles, out: count * 5.
hipError_t launch_reduce_boxes(const lbm_solver* const* s, const Box* b, int count,
                               double* partials, double* out_dev, hipStream_t st);

// Default launch options of the step kernels (kernel 0 = alpha, 1 = beta),
// measured on MI355X: FP32 and FP64 gain from fewer resident waves (fewer
// concurrent plane streams), 16-bit storage needs every wave; F64 beta from
// line-aligned accesses staged through LDS. FP32 from 512-thread workgroups
// of two cells per thread on large domains (wide: at 192^3 and up; below
// about 160^3 they leave too few workgroups); with the fused x-face pack
// (with_pack) FP32 beta on large domains runs in 1024-thread workgroups, one
// per CU, and FP32 alpha on small domains keeps 3 waves per SIMD (measured
// with the six-face exchange at 512^3). 16-bit storage: cached
// accesses; F16S and BF16S beta in 1024-thread workgroups on large domains
// (F16C beta, with more integer work per cell, loses with them).
constexpr int64_t kWideCells = 6 << 20;   // halo-inclusive cells for `wide`

// 32-bit byte offsets reach every element of a plane
inline bool plane_off32(const lbm_solver* s)
{
    return (uint64_t)s->stride * (uint64_t)s->elem_bytes < (1ull << 32);
}

// addr64: the planes need 64-bit element offsets (4 GiB and up); the FP32
// beta with the fused unpack in 1024-thread workgroups then runs out of its
// 128 registers per thread and spills (half the step rate at 512^3 forced to
// 64-bit offsets), so it keeps 512-thread workgroups.
inline int default_launch(int precision, int kernel, bool with_pack, bool wide,
                          bool with_unpack = false, bool addr64 = false)
{
    switch (precision) {
    case LBM_F32:
        if (kernel == 0) {
            // 256-thread workgroups, 3 waves per SIMD: with a step pack on a
            // small domain, and with the fused x-face unpack (whose received
            // loads hold a wave longer than the streaming 2-wave shape allows)
            if (with_unpack || (with_pack && !wide))
                return LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_WAVES3;
            return (wide ? LBM_LAUNCH_WIDE : 0) | LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_WAVES2;
        }
        if (with_unpack && wide && addr64)
            return LBM_LAUNCH_WIDE | LBM_LAUNCH_WAVES3;
        if ((with_pack || with_unpack) && wide)
            return LBM_LAUNCH_BIG | LBM_LAUNCH_WAVES4;
        return (wide ? LBM_LAUNCH_WIDE : 0) | LBM_LAUNCH_WAVES2;
    case LBM_F64:
        return kernel == 0 ? (LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_WAVES2)
                           : (LBM_LAUNCH_LDS | LBM_LAUNCH_WAVES2);
    case LBM_F16C:
        return 0;
    default:
        // 1024-thread beta workgroups wait on their slowest wave: with the
        // fused unpack's received loads 256-thread ones are faster
        return kernel == 1 && wide && !with_unpack ? LBM_LAUNCH_BIG : 0;
    }
}

// ||u - u_prev|| / ||u|| from the two FP64 sums; a non-finite state gives NaN
// (never a small number that would pass a convergence test).
inline double residual_ratio(double num, double den)
{
    if (std::isnan(num) || std::isnan(den) || std::isinf(num) || std::isinf(den))
        return std::numeric_limits<double>::quiet_NaN();
    if (den > 0)
        return std::sqrt(num / den);
    return num > 0 ? std::numeric_limits<double>::infinity() : 0.0;
}

// The residual call's value: residual_ratio, except on the first call after
// creation, reset or load, where no previous field exists (num compares with
// a stale snapshot): +inf, or NaN when the current field (den) is already
// non-finite.
inline double first_residual(int valid, double num, double den)
{
    if (valid)
        return residual_ratio(num, den);
    if (std::isnan(den) || std::isinf(den))
        return std::numeric_limits<double>::quiet_NaN();
    return std::numeric_limits<double>::infinity();
}

}  // namespace lbm
