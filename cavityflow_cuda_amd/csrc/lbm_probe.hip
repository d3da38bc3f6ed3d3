// Bandwidth probes of the step kernels' memory access patterns, without the
// collision: the ceilings that bound k_alpha and k_beta on this memory system
// (bench.py --workload ceiling). They run on a solver's own population array
// (19 planes at the solver's plane stride, elements of its storage size: 4
// bytes FP32, 8 FP64, 2 for the 16-bit formats) and leave its values
// unchanged, except the write pattern, which stores the rest state.
//
//   RMW_ALIGNED : each cell loads and stores its own 19 slots (alpha).
//   RMW_SHIFTED : each cell loads and stores plane l at id + c_l (beta), for
//                 the cells whose 19 neighbours lie inside the array.
//   RMW_SHIFTED_PITCHED: beta's pattern as a layout with rows padded to a
//                 multiple of 32 elements would give it: shifts
//                 c_x + c_y P + c_z P ny, P = nx rounded up to 32 (y and z
//                 shifts line-aligned for FP32, x shifts by one element).
//   RMW_SHIFTED_X: the x components of the shifts only (one element).
//   RMW_SHIFTED_YZ: the y and z components only (nx and nx ny elements).
//   RMW_LANE_SHIFTED (FP32): the padded-row shifts with the x component
//                 taken across lanes: every wave loads and stores whole
//                 aligned 64-cell chunks of a plane, moves the values one
//                 lane up or down (DPP wave shifts), and adds one element
//                 beyond either end of its run of consecutive chunks.
//   READ        : 19 loads per cell, nothing stored.
//   WRITE       : 19 stores per cell.
// Launch options: LBM_LAUNCH_NONTEMPORAL and the WAVES4/3/2 occupancy caps,
// as for the step kernels; the workgroup shape of the read-modify-write
// patterns (LBM_PROBE_BLOCK512 / 1024 threads, LBM_PROBE_SPAN2 / 4 / 8 cells
// per thread one block apart, as the step kernels' WIDE shape does with 2):
// each workgroup then covers block x span consecutive cells; or, FP32 only,
// LBM_PROBE_VEC2 / VEC4: 2 or 4 consecutive cells per thread, one vector
// access per plane.
#include "lbm_d3q19.hpp"
#include "lbm_internal.hpp"

#include <algorithm>

namespace lbm {

namespace {

struct ProbeArgs {
    const char* plane[kQ];   // byte base of plane l (RMW_SHIFTED: shifted by c_l)
    uint32_t first, last;    // cells [first, last)
    float zero;              // 0.0f at run time: keeps the stored value a function of the load
    uint32_t zero_bits;      // 0 at run time, the same for 16-bit codes
    uint32_t never;          // 0xFFFFFFFF: no sum of 19 16-bit codes reaches it
};

// The stored value: the loaded one through a run-time identity.
__device__ inline float keep(float v, const ProbeArgs& a) { return v + a.zero; }
__device__ inline double keep(double v, const ProbeArgs& a) { return v + (double)a.zero; }
__device__ inline uint16_t keep(uint16_t v, const ProbeArgs& a)
{
    return (uint16_t)(v ^ a.zero_bits);
}

// The rest state of plane l: the lattice weight, or code 0 (f - w = 0) for
// the 16-bit formats.
template <class E>
__device__ inline E rest(int l)
{
    if constexpr (sizeof(E) == 2)
        return 0;
    else
        return weight<E>(l);
}

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
__global__ __launch_bounds__(256) void k_probe_vec(ProbeArgs a)
{
    using V = typename VecOf<VEC>::T;
    const uint32_t id = a.first + (blockIdx.x * 256 + threadIdx.x) * VEC;
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

// One cell: its 19 accesses at the shared byte offset of plane l's base.
template <class E, int MODE, bool NT>
__device__ inline void probe_cell(const ProbeArgs& a, uint32_t id)
{
    const uint32_t off = id * (uint32_t)sizeof(E);
    E v[kQ];
    if (MODE != LBM_PROBE_WRITE) {
#pragma unroll
        for (int l = 0; l < kQ; l++) {
            const __attribute__((address_space(1))) E* p =
                (const __attribute__((address_space(1))) E*)(a.plane[l] + off);
            v[l] = NT ? __builtin_nontemporal_load(p) : *p;
        }
    }
    if (MODE == LBM_PROBE_READ) {
        // a store that never happens keeps the loads live
        if constexpr (sizeof(E) == 2) {
            uint32_t s = 0;
#pragma unroll
            for (int l = 0; l < kQ; l++)
                s += v[l];
            if (s == a.never)
                *(__attribute__((address_space(1))) E*)(const_cast<char*>(a.plane[0]) + off) =
                    (E)s;
        } else {
            E s = 0;
#pragma unroll
            for (int l = 0; l < kQ; l++)
                s += v[l];
            if (s == (E)1e30f)   // populations sum to the density
                *(__attribute__((address_space(1))) E*)(const_cast<char*>(a.plane[0]) + off) = s;
        }
        return;
    }
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        const E w = MODE == LBM_PROBE_WRITE ? keep(rest<E>(l), a) : keep(v[l], a);
        uint32_t o = off;
        asm volatile("" : "+v"(o));
        __attribute__((address_space(1))) E* p =
            (__attribute__((address_space(1))) E*)(const_cast<char*>(a.plane[l]) + o);
        if (NT)
            __builtin_nontemporal_store(w, p);
        else
            *p = w;
    }
}

// BLOCK threads per workgroup, SPAN cells per thread one block apart: the
// workgroup covers BLOCK * SPAN consecutive cells.
template <class E, int MODE, bool NT, int BLOCK, int SPAN>
__global__ __launch_bounds__(BLOCK) void k_probe(ProbeArgs a)
{
    const uint32_t base = a.first + blockIdx.x * (uint32_t)(BLOCK * SPAN) + threadIdx.x;
#pragma unroll 1
    for (int k = 0; k < SPAN; k++) {
        const uint32_t id = base + k * BLOCK;
        if (id >= a.last)
            return;
        probe_cell<E, MODE, NT>(a, id);
    }
}

// RMW_LANE_SHIFTED: wave w of a workgroup takes R consecutive 64-cell chunks
// from a line-aligned start; plane l's base carries the (line-aligned) y / z
// shift. For c_x = +1 cell i takes slot i + 1: lane j gets lane j + 1 of its
// chunk (wave_shl:1), lane 63 lane 0 of the next chunk, the last chunk's lane
// 63 the element past the run, loaded on its own; c_x = -1 mirrors it. The
// stores move the values back: each chunk's aligned slots, less the one slot
// a neighbouring wave owns, plus that end element as a single-lane store.
__device__ inline float lane_from(float old, float v, int ctrl)
{
    // ctrl is 0x130 (wave_shl:1, lane j reads lane j + 1) or 0x138 (wave_shr:1,
    // lane j reads lane j - 1); the lane with no source keeps old
    return __int_as_float(ctrl == 0x130
        ? __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x130, 0xf, 0xf, false)
        : __builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), 0x138, 0xf, 0xf, false));
}

__device__ inline float lane_value(float v, int lane)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

template <bool NT>
__device__ inline float ld(const char* base, uint32_t cell)
{
    const __attribute__((address_space(1))) float* p =
        (const __attribute__((address_space(1))) float*)(base + cell * (uint32_t)sizeof(float));
    return NT ? __builtin_nontemporal_load(p) : *p;
}

template <bool NT>
__device__ inline void st(const char* base, uint32_t cell, float v)
{
    uint32_t o = cell * (uint32_t)sizeof(float);
    asm volatile("" : "+v"(o));
    __attribute__((address_space(1))) float* p =
        (__attribute__((address_space(1))) float*)(const_cast<char*>(base) + o);
    if (NT)
        __builtin_nontemporal_store(v, p);
    else
        *p = v;
}

template <bool NT, int BLOCK, int R>
__global__ __launch_bounds__(BLOCK) void k_probe_lanes(ProbeArgs a)
{
    constexpr int kShl = 0x130, kShr = 0x138;
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t w0 =
        a.first + (blockIdx.x * (uint32_t)(BLOCK / 64) + (threadIdx.x >> 6)) * (uint32_t)(64 * R);
    if (w0 >= a.last)
        return;   // [first, last) is whole chunks: every wave below has 1..R of them
    const int nch = (int)min((a.last - w0) / 64u, (uint32_t)R);
    float v[kQ][R];
#pragma unroll
    for (int l = 0; l < kQ; l++) {
#pragma unroll
        for (int k = 0; k < R; k++)
            v[l][k] = k < nch ? ld<NT>(a.plane[l], w0 + 64 * k + lane) : 0.f;
    }
    float e[kQ];
#pragma unroll
    for (int l = 0; l < kQ; l++)
        e[l] = cx(l) > 0 ? ld<NT>(a.plane[l], w0 + 64 * nch)
               : cx(l) < 0 ? ld<NT>(a.plane[l], w0 - 1) : 0.f;
    // the cells' values, as a step would take them
#pragma unroll
    for (int l = 0; l < kQ; l++) {
        float c[R];
#pragma unroll
        for (int k = 0; k < R; k++) {
            if (k >= nch)
                break;
            if (cx(l) > 0)
                c[k] = lane_from(k + 1 < R && k + 1 < nch ? lane_value(v[l][k + 1 < R ? k + 1 : k], 0) : e[l],
                                 v[l][k], kShl);
            else if (cx(l) < 0)
                c[k] = lane_from(k > 0 ? lane_value(v[l][k > 0 ? k - 1 : 0], 63) : e[l], v[l][k], kShr);
            else
                c[k] = v[l][k];
            c[k] = keep(c[k], a);
        }
        // back to the slots they came from. Every cross-lane read comes
        // before the lane-dependent stores: inside a branch that leaves lane 63
        // inactive the compiler may reuse that lane's registers, and a
        // readlane there returns whatever they then hold.
#pragma unroll
        for (int k = 0; k < R; k++) {
            if (k >= nch)
                break;
            const uint32_t cell = w0 + 64 * k + lane;
            if (cx(l) > 0) {
                const float w = lane_from(k > 0 ? lane_value(c[k > 0 ? k - 1 : 0], 63) : 0.f, c[k], kShr);
                const float end = lane_value(c[k], 63);
                if (k > 0 || lane > 0)
                    st<NT>(a.plane[l], cell, w);
                if (k == nch - 1 && lane == 0)
                    st<NT>(a.plane[l], w0 + 64 * nch, end);
            } else if (cx(l) < 0) {
                const float w = lane_from(k + 1 < R && k + 1 < nch ? lane_value(c[k + 1 < R ? k + 1 : k], 0) : 0.f,
                                          c[k], kShl);
                const float end = lane_value(c[k], 0);
                if (k < nch - 1 || lane < 63)
                    st<NT>(a.plane[l], cell, w);
                if (k == 0 && lane == 0)
                    st<NT>(a.plane[l], w0 - 1, end);
            } else {
                st<NT>(a.plane[l], cell, c[k]);
            }
        }
    }
}

// Dynamic LDS that caps a CU at m resident workgroups (160 KB of LDS per CU).
size_t cap_lds(int m)
{
    return m >= 4 ? 33 * 1024 : m == 3 ? 41 * 1024 : m == 2 ? 54 * 1024 : 81 * 1024;
}

template <int MODE, int VEC>
hipError_t launch_vec(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    const uint32_t threads = (a.last - a.first) / VEC;
    const uint32_t blocks = (threads + 255) / 256;
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_probe_vec<MODE, true, VEC>), dim3(blocks), dim3(256), lds, st, a);
    else
        hipLaunchKernelGGL((k_probe_vec<MODE, false, VEC>), dim3(blocks), dim3(256), lds, st, a);
    return hipGetLastError();
}

template <int BLOCK, int R>
hipError_t launch_lanes_r(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    const uint32_t chunks = (a.last - a.first) / 64;
    const uint32_t per = (uint32_t)(BLOCK / 64 * R);
    const uint32_t blocks = (chunks + per - 1) / per;
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_probe_lanes<true, BLOCK, R>), dim3(blocks), dim3(BLOCK), lds, st, a);
    else
        hipLaunchKernelGGL((k_probe_lanes<false, BLOCK, R>), dim3(blocks), dim3(BLOCK), lds, st,
                           a);
    return hipGetLastError();
}

template <int BLOCK>
hipError_t launch_lanes(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    if constexpr (BLOCK < 1024)
        if (opts & LBM_PROBE_SPAN8)
            return launch_lanes_r<BLOCK, 8>(a, opts, lds, st);
    if (opts & LBM_PROBE_SPAN4)
        return launch_lanes_r<BLOCK, 4>(a, opts, lds, st);
    if (opts & LBM_PROBE_SPAN2)
        return launch_lanes_r<BLOCK, 2>(a, opts, lds, st);
    return launch_lanes_r<BLOCK, 1>(a, opts, lds, st);
}

template <class E, int MODE, int BLOCK, int SPAN>
hipError_t launch_span(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    const uint32_t blocks = (a.last - a.first + BLOCK * SPAN - 1) / (BLOCK * SPAN);
    if (opts & LBM_LAUNCH_NONTEMPORAL)
        hipLaunchKernelGGL((k_probe<E, MODE, true, BLOCK, SPAN>), dim3(blocks), dim3(BLOCK), lds,
                           st, a);
    else
        hipLaunchKernelGGL((k_probe<E, MODE, false, BLOCK, SPAN>), dim3(blocks), dim3(BLOCK),
                           lds, st, a);
    return hipGetLastError();
}

template <class E, int MODE, int BLOCK>
hipError_t launch_block(const ProbeArgs& a, int opts, size_t lds, hipStream_t st)
{
    if (MODE == LBM_PROBE_READ || MODE == LBM_PROBE_WRITE)
        return launch_span<E, MODE, BLOCK, 1>(a, opts, lds, st);
    if (opts & LBM_PROBE_SPAN8)
        return launch_span<E, MODE, BLOCK, 8>(a, opts, lds, st);
    if (opts & LBM_PROBE_SPAN4)
        return launch_span<E, MODE, BLOCK, 4>(a, opts, lds, st);
    if (opts & LBM_PROBE_SPAN2)
        return launch_span<E, MODE, BLOCK, 2>(a, opts, lds, st);
    return launch_span<E, MODE, BLOCK, 1>(a, opts, lds, st);
}

template <class E, int MODE>
hipError_t launch_mode(const ProbeArgs& a, int opts, hipStream_t st)
{
    const int block = (opts & LBM_PROBE_BLOCK1024) ? 1024 : (opts & LBM_PROBE_BLOCK512) ? 512 : 256;
    const int waves = (opts & LBM_LAUNCH_WAVES2) ? 2 : (opts & LBM_LAUNCH_WAVES3) ? 3
                      : (opts & LBM_LAUNCH_WAVES4) ? 4 : 0;
    // a cap of w waves per SIMD leaves w * 256 / block workgroups per CU
    const size_t lds = waves ? cap_lds(waves * 256 / block) : 0;
    if constexpr (sizeof(E) == 4) {
        if (MODE == LBM_PROBE_RMW_LANE_SHIFTED)
            return block == 1024 ? launch_lanes<1024>(a, opts, lds, st)
                   : block == 512 ? launch_lanes<512>(a, opts, lds, st)
                                  : launch_lanes<256>(a, opts, lds, st);
        if (MODE != LBM_PROBE_READ && MODE != LBM_PROBE_WRITE) {
            if (opts & LBM_PROBE_VEC4)
                return launch_vec<MODE, 4>(a, opts, lds, st);
            if (opts & LBM_PROBE_VEC2)
                return launch_vec<MODE, 2>(a, opts, lds, st);
        }
    }
    if (block == 1024)
        return launch_block<E, MODE, 1024>(a, opts, lds, st);
    if (block == 512)
        return launch_block<E, MODE, 512>(a, opts, lds, st);
    return launch_block<E, MODE, 256>(a, opts, lds, st);
}

template <class E>
hipError_t launch_probe(const ProbeArgs& a, int mode, int opts, hipStream_t st)
{
    switch (mode) {
    case LBM_PROBE_RMW_ALIGNED: return launch_mode<E, LBM_PROBE_RMW_ALIGNED>(a, opts, st);
    case LBM_PROBE_READ: return launch_mode<E, LBM_PROBE_READ>(a, opts, st);
    case LBM_PROBE_WRITE: return launch_mode<E, LBM_PROBE_WRITE>(a, opts, st);
    case LBM_PROBE_RMW_LANE_SHIFTED:
        return launch_mode<E, LBM_PROBE_RMW_LANE_SHIFTED>(a, opts, st);
    default:   // the shifted variants differ only in their plane bases
        return launch_mode<E, LBM_PROBE_RMW_SHIFTED>(a, opts, st);
    }
}

}  // namespace

}  // namespace lbm

using namespace lbm;

extern "C" int lbm_probe_pattern(lbm_solver* s, int mode, int options, uint64_t* cells,
                                 hipStream_t stream)
{
    const int known = LBM_LAUNCH_NONTEMPORAL | LBM_LAUNCH_WAVES4 | LBM_LAUNCH_WAVES3 |
                      LBM_LAUNCH_WAVES2 | LBM_PROBE_VEC2 | LBM_PROBE_VEC4 | LBM_PROBE_BLOCK512 |
                      LBM_PROBE_BLOCK1024 | LBM_PROBE_SPAN2 | LBM_PROBE_SPAN4 | LBM_PROBE_SPAN8;
    if (!s || mode < LBM_PROBE_RMW_ALIGNED || mode > LBM_PROBE_RMW_LANE_SHIFTED ||
        (options & ~known)) {
        set_error("lbm_probe_pattern: invalid argument");
        return LBM_ERR_ARG;
    }
    const bool rmw = mode != LBM_PROBE_READ && mode != LBM_PROBE_WRITE;
    const int vec = options & (LBM_PROBE_VEC2 | LBM_PROBE_VEC4);
    const int shape = options & (LBM_PROBE_BLOCK512 | LBM_PROBE_BLOCK1024 | LBM_PROBE_SPAN2 |
                                 LBM_PROBE_SPAN4 | LBM_PROBE_SPAN8);
    const auto one_bit = [](int v) { return (v & (v - 1)) == 0; };
    const int spans = options & (LBM_PROBE_SPAN2 | LBM_PROBE_SPAN4 | LBM_PROBE_SPAN8);
    const int blocks = options & (LBM_PROBE_BLOCK512 | LBM_PROBE_BLOCK1024);
    if (!one_bit(vec) || !one_bit(spans) || !one_bit(blocks) || (vec && shape) ||
        (!rmw && (vec || shape))) {
        set_error("lbm_probe_pattern: VEC2 / VEC4, BLOCK512 / BLOCK1024 and SPAN2 / 4 / 8 "
                  "each pick one, vector accesses exclude the workgroup shapes, and the "
                  "read / write streams take neither");
        return LBM_ERR_ARG;
    }
    const int block = (options & LBM_PROBE_BLOCK1024) ? 1024 : (options & LBM_PROBE_BLOCK512) ? 512
                                                                                              : 256;
    const int waves = (options & LBM_LAUNCH_WAVES2) ? 2 : (options & LBM_LAUNCH_WAVES3) ? 3
                      : (options & LBM_LAUNCH_WAVES4) ? 4 : 0;
    if (waves && waves * 256 < block) {
        set_error("lbm_probe_pattern: that waves-per-SIMD cap holds less than one "
                  "workgroup of this size");
        return LBM_ERR_ARG;
    }
    const size_t elem = (size_t)s->elem_bytes;
    const bool lanes = mode == LBM_PROBE_RMW_LANE_SHIFTED;
    if (lanes && (vec || ((options & LBM_PROBE_SPAN8) && block == 1024))) {
        set_error("lbm_probe_pattern: the lane-shifted pattern takes no vector accesses, "
                  "and 1024-thread workgroups at most 4 chunks per wave");
        return LBM_ERR_ARG;
    }
    if ((vec || lanes) && s->p.precision != LBM_F32) {
        set_error("lbm_probe_pattern: vector accesses and the lane-shifted pattern on FP32 "
                  "solvers only");
        return LBM_ERR_UNSUPPORTED;
    }
    if ((uint64_t)s->stride * elem >= (1ull << 32)) {
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
        // the x shift is taken across lanes; one more element at either end
        case LBM_PROBE_RMW_LANE_SHIFTED: shift[l] = cy(l) * pitch + cz(l) * pxy; break;
        default: shift[l] = 0;
        }
        span = std::max<int64_t>(span, shift[l] < 0 ? -shift[l] : shift[l]);
    }
    if (lanes)
        span += 1;
    // every accessed element id + shift inside [0, n); the lane-shifted
    // pattern covers whole 64-cell chunks from a 64-cell boundary
    int64_t first = span, last = s->n - span;
    if (lanes) {
        first = (first + 63) / 64 * 64;
        last = last < first ? first : first + (last - first) / 64 * 64;
    }
    if (2 * span >= s->n || last <= first) {
        set_error("lbm_probe_pattern: domain too small for this pattern");
        return LBM_ERR_ARG;
    }
    a.first = (uint32_t)first;
    a.last = (uint32_t)last;
    a.zero = 0.0f;
    a.zero_bits = 0;
    a.never = 0xFFFFFFFFu;
    for (int l = 0; l < kQ; l++)
        a.plane[l] = (const char*)s->dd + ((int64_t)l * s->stride + shift[l]) * (int64_t)elem;
    const uint32_t per = (options & LBM_PROBE_VEC4) ? 4 : (options & LBM_PROBE_VEC2) ? 2 : 1;
    if (cells)
        *cells = (a.last - a.first) / per * per;
    hipError_t e;
    switch (elem) {
    case 8: e = launch_probe<double>(a, mode, options, stream); break;
    case 2: e = launch_probe<uint16_t>(a, mode, options, stream); break;
    default: e = launch_probe<float>(a, mode, options, stream); break;
    }
    if (e != hipSuccess)
        return hip_fail(e, "probe kernel launch");
    return LBM_OK;
}
