// Bandwidth probes of the step kernels' memory access patterns, without the
// collision: the ceilings that bound k_alpha and k_beta on this memory system
// (bench.py --workload ceiling). They run on a solver's own population array
// (19 planes at the solver's plane stride, one thread per cell, 256-thread
// workgroups) and leave its values unchanged, except the write pattern, which
// stores the lattice weights (the rest state).
//
//   RMW_ALIGNED : each cell loads and stores its own 19 slots (alpha).
//   RMW_SHIFTED : each cell loads and stores plane l at id + c_l (beta), for
//                 the cells whose 19 neighbours lie inside the array.
//   RMW_SHIFTED_PITCHED: beta's pattern as a layout with rows padded to a
//                 multiple of 32 elements (128 bytes) would give it: shifts
//                 c_x + c_y P + c_z P ny, P = nx rounded up to 32 (y and z
//                 shifts line-aligned, x shifts by one element).
//   RMW_SHIFTED_X: the x components of the shifts only (one element).
//   RMW_SHIFTED_YZ: the y and z components only (nx and nx ny elements).
//   READ        : 19 loads per cell, nothing stored.
//   WRITE       : 19 stores per cell.
// Launch options: LBM_LAUNCH_NONTEMPORAL and the WAVES4/3/2 occupancy caps,
// as for the step kernels; LBM_PROBE_VEC2 / VEC4 give the read-modify-write
// patterns 2 or 4 consecutive cells per thread, one vector access per plane.
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"

#include <algorithm>

namespace lbm {

namespace {

constexpr int kProbeBlock = 256;

struct ProbeArgs {
    const char* plane[kQ];   // byte base of plane l (RMW_SHIFTED: shifted by c_l)
    uint32_t first, last;    // cells [first, last)
    float zero;              // 0.0f at run time: keeps the stored value a function of the load
};

// VEC consecutive cells per thread, accessed as one VEC-float vector per
// plane (4-byte aligned: the shifted planes start at any element).
template <int VEC>
struct VecOf;
template <>
struct VecOf<2> {
    typedef float T __attribute__((ext_vector_type(2), aligned(4)));
};
template <>
struct VecOf<4> {
    typedef float T __attribute__((ext_vector_type(4), aligned(4)));
};

template <int MODE, bool NT, int VEC>
__global__ __launch_bounds__(kProbeBlock) void k_probe_vec(ProbeArgs a)
{
    using V = typename VecOf<VEC>::T;
    const uint32_t id = a.first + (blockIdx.x * kProbeBlock + threadIdx.x) * VEC;
    if (id + VEC > a.last)
        return;
    const uint32_t off = id * (uint32_t)sizeof(float);
    V v[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const __attribute__((address_space(1))) V* p =
            (const __attribute__((address_space(1))) V*)(a.plane[l] + off);
        v[l] = NT ? __builtin_nontemporal_load(p) : *p;
    }
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const V w = v[l] + a.zero;
        uint32_t o = off;
        asm volatile("" : "+v"(o));
        __attribute__((address_space(1))) V* p =
            (__attribute__((address_space(1))) V*)(const_cast<char*>(a.plane[l]) + o);
        if (NT)
            __builtin_nontemporal_store(w, p);
        else
            *p = w;
    }
}

template <int MODE, bool NT>
__global__ __launch_bounds__(kProbeBlock) void k_probe(ProbeArgs a)
{
    const uint32_t id = a.first + blockIdx.x * kProbeBlock + threadIdx.x;
    if (id >= a.last)
        return;
    const uint32_t off = id * (uint32_t)sizeof(float);
    float v[kQ];
    if (MODE != LBM_PROBE_WRITE) {
#pragma unroll
        for (int l = 0; l < kQ; l++) {
            const __attribute__((address_space(1))) float* p =
                (const __attribute__((address_space(1))) float*)(a.plane[l] + off);
            v[l] = NT ? __builtin_nontemporal_load(p) : *p;
        }
    }
    if (MODE == LBM_PROBE_READ) {
        float s = 0.f;
#pragma unroll
        for (int l = 0; l < kQ; l++)
            s += v[l];
        if (s == 1e30f)   // never: the loads stay live, nothing is stored
            *(__attribute__((address_space(1))) float*)(const_cast<char*>(a.plane[0]) + off) = s;
        return;
    }
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const float w = MODE == LBM_PROBE_WRITE ? weight<float>(l) + a.zero : v[l] + a.zero;
        uint32_t o = off;
        asm volatile("" : "+v"(o));
        __attribute__((address_space(1))) float* p =
            (__attribute__((address_space(1))) float*)(const_cast<char*>(a.plane[l]) + o);
        if (NT)
            __builtin_nontemporal_store(w, p);
        else
            *p = w;
    }
}

template <int MODE, int VEC>
hipError_t launch_vec(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    const uint32_t threads = (a.last - a.first) / VEC;
    const uint32_t blocks = (threads + kProbeBlock - 1) / kProbeBlock;
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_probe_vec<MODE, true, VEC>), dim3(blocks), dim3(kProbeBlock), lds,
                           st, a);
    else
        hipLaunchKernelGGL((k_probe_vec<MODE, false, VEC>), dim3(blocks), dim3(kProbeBlock), lds,
                           st, a);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_mode(const ProbeArgs& a, int opts, hipStream_t st)
{
    const size_t lds = (opts & LBM_LAUNCH_WAVES2)   ? 54 * 1024
                       : (opts & LBM_LAUNCH_WAVES3) ? 41 * 1024
                       : (opts & LBM_LAUNCH_WAVES4) ? 33 * 1024 : 0;
    if (MODE != LBM_PROBE_READ && MODE != LBM_PROBE_WRITE) {
        if (opts & LBM_PROBE_VEC4)
            return launch_vec<MODE, 4>(a, opts, lds, st);
        if (opts & LBM_PROBE_VEC2)
            return launch_vec<MODE, 2>(a, opts, lds, st);
    }
    const uint32_t blocks = (a.last - a.first + kProbeBlock - 1) / kProbeBlock;
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_probe<MODE, true>), dim3(blocks), dim3(kProbeBlock), lds, st, a);
    else
        hipLaunchKernelGGL((k_probe<MODE, false>), dim3(blocks), dim3(kProbeBlock), lds, st, a);
    return hipGetLastError();
}

}  // namespace

}  // namespace lbm

using namespace lbm;

extern "C" int lbm_probe_pattern(lbm_solver* s, int mode, int options, uint64_t* cells,
                                 hipStream_t stream)
{
    if (!s || mode < LBM_PROBE_RMW_ALIGNED || mode > LBM_PROBE_RMW_SHIFTED_YZ ||
        (options & ~(LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_WAVES4 | LBM_LAUNCH_WAVES3 |
                     LBM_LAUNCH_WAVES2 | LBM_PROBE_VEC2 | LBM_PROBE_VEC4))) {
        set_error("lbm_probe_pattern: invalid argument");
        return LBM_ERR_ARG;
    }
    if (s->p.precision != LBM_F32) {
        set_error("lbm_probe_pattern: FP32 solvers only");
        return LBM_ERR_UNSUPPORTED;
    }
    if ((uint64_t)s->stride * sizeof(float) >= (1ull << 32)) {
        set_error("lbm_probe_pattern: planes beyond 32-bit offsets");
        return LBM_ERR_UNSUPPORTED;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || (cur != s->p.device && hipSetDevice(s->p.device)))
        return hip_fail(hipErrorInvalidDevice, "hipSetDevice");
    ProbeArgs a;
    const int64_t nx = s->size_h[0], nxy = nx * s->size_h[1];
    const int64_t pitch = (nx + 31) / 32 * 32, pxy = pitch * s->size_h[1];
    int64_t shift[kQ], span = 0;
    for (int l = 0; l < kQ; l++) {
        switch (mode) {
        case LBM_PROBE_RMW_SHIFTED: shift[l] = cx(l) + cy(l) * nx + cz(l) * nxy; break;
        case LBM_PROBE_RMW_SHIFTED_PITCHED: shift[l] = cx(l) + cy(l) * pitch + cz(l) * pxy; break;
        case LBM_PROBE_RMW_SHIFTED_X: shift[l] = cx(l); break;
        case LBM_PROBE_RMW_SHIFTED_YZ: shift[l] = cy(l) * nx + cz(l) * nxy; break;
        default: shift[l] = 0;
        }
        span = std::max<int64_t>(span, shift[l] < 0 ? -shift[l] : shift[l]);
    }
    if (2 * span >= s->n) {
        set_error("lbm_probe_pattern: domain too small for this pattern");
        return LBM_ERR_ARG;
    }
    // every accessed element id + shift inside [0, n)
    a.first = (uint32_t)span;
    a.last = (uint32_t)(s->n - span);
    a.zero = 0.0f;
    for (int l = 0; l < kQ; l++)
        a.plane[l] = (const char*)s->dd +
                     ((int64_t)l * s->stride + shift[l]) * (int64_t)sizeof(float);
    const bool vec = mode != LBM_PROBE_READ && mode != LBM_PROBE_WRITE;
    const uint32_t per = vec ? ((options & LBM_PROBE_VEC4) ? 4 : (options & LBM_PROBE_VEC2) ? 2 : 1)
                             : 1;
    if (cells)
        *cells = (a.last - a.first) / per * per;
    hipError_t e = hipErrorInvalidValue;
    switch (mode) {
    case LBM_PROBE_RMW_ALIGNED: e = launch_mode<LBM_PROBE_RMW_ALIGNED>(a, options, stream); break;
    case LBM_PROBE_RMW_SHIFTED: e = launch_mode<LBM_PROBE_RMW_SHIFTED>(a, options, stream); break;
    case LBM_PROBE_READ: e = launch_mode<LBM_PROBE_READ>(a, options, stream); break;
    case LBM_PROBE_WRITE: e = launch_mode<LBM_PROBE_WRITE>(a, options, stream); break;
    default:   // the shifted variants differ only in their plane bases
        e = launch_mode<LBM_PROBE_RMW_SHIFTED>(a, options, stream);
        break;
    }
    if (e != hipSuccess)
        return hip_fail(e, "probe kernel launch");
    return LBM_OK;
}
