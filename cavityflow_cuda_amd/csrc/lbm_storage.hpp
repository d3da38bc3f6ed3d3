// Population storage policies of the HIP solver.
//
//   F32   : float populations, float arithmetic (bit-exact with the reference)
//   F64   : double populations, double arithmetic (bit-exact; FP64 oracle runs)
//   F16S  : IEEE binary16 of (f_l - w_l) * 2^12, float arithmetic
//   BF16S : bfloat16 of (f_l - w_l), float arithmetic
//
// The shifted forms store the deviation from the lattice weight w_l, which
// keeps the stored magnitude near 1e-3 where 16-bit formats still hold ~3
// significant digits. The weights are symmetric (w_l = w_opp(l)), so plane l
// uses w_l in both the alpha (swapped) and beta (natural) layouts.
#pragma once

#include <hip/hip_fp16.h>
#include <stdint.h>

#include "lbm_d3q19.hpp"

namespace lbm {

struct StoreF32 {
    using S = float;
    using C = float;
    static constexpr bool kShifted = false;
    __device__ static __forceinline__ C load(S v, int) { return v; }
    __device__ static __forceinline__ S store(C v, int) { return v; }
};

struct StoreF64 {
    using S = double;
    using C = double;
    static constexpr bool kShifted = false;
    __device__ static __forceinline__ C load(S v, int) { return v; }
    __device__ static __forceinline__ S store(C v, int) { return v; }
};

// FP16 keeps 11 significant bits only above 2^-14; slow regions of a cavity
// have |f - w| ~ 1e-5, which would fall into its subnormal range. The stored
// value is (f - w) * 2^12: exact scaling, normal range down to |f - w| ~ 1.5e-8,
// and |f - w| <= ~0.1 stays far below the FP16 maximum (65504).
struct StoreF16S {
    using S = _Float16;
    using C = float;
    static constexpr bool kShifted = true;
    static constexpr float kScale = 4096.0f;
    // (float)v * 2^-12 is exact, so the fused multiply-add rounds once, exactly
    // as the separate multiply and add do; it compiles to one mixed-precision
    // FMA that also converts v
    __device__ static __forceinline__ C load(S v, int l)
    {
        return __builtin_fmaf((float)v, 1.0f / kScale, weight<float>(l));
    }
    __device__ static __forceinline__ S store(C v, int l)
    {
        return (_Float16)((v - weight<float>(l)) * kScale);
    }
};

struct StoreBF16S {
    using S = uint16_t;  // raw bfloat16 bits
    using C = float;
    static constexpr bool kShifted = true;
    __device__ static __forceinline__ float bf16_to_float(uint16_t b)
    {
        return __uint_as_float(((uint32_t)b) << 16);
    }
    // round to nearest even (gfx950 converts in one instruction); inputs are
    // finite
    __device__ static __forceinline__ uint16_t float_to_bf16(float f)
    {
        return __builtin_bit_cast(uint16_t, (__bf16)f);
    }
    __device__ static __forceinline__ C load(S v, int l) { return bf16_to_float(v) + weight<float>(l); }
    __device__ static __forceinline__ S store(C v, int l) { return float_to_bf16(v - weight<float>(l)); }
};

}  // namespace lbm
