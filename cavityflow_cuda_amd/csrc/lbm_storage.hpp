// Population storage policies of the HIP solver.
//
//   F32   : float populations, float arithmetic (bit-exact with the reference)
//   F64   : double populations, double arithmetic (bit-exact; FP64 oracle runs)
//   F16S  : IEEE binary16 of (f_l - w_l) * 2^12, float arithmetic
//   BF16S : bfloat16 of (f_l - w_l), float arithmetic
//   F16C  : 16-bit float of (f_l - w_l) with 4 exponent and 11 significand
//           bits over |f - w| < 1/4, float arithmetic
//
// Every policy has load (stored value -> arithmetic value of plane l) and
// store (the inverse); store_checked also ORs a range word into `ov`, and
// overflowed(ov) says whether any checked store was out of range (only F16C,
// the one format without infinities, can overflow without a trace).
// The shifted forms store the deviation from the lattice weight w_l, which
// keeps the stored magnitude near 1e-3 where 16-bit formats still hold ~3
// significant digits. The weights are symmetric (w_l = w_opp(l)), so plane l
// uses w_l in both the alpha (swapped) and beta (natural) layouts.
#pragma once

#include <hip/hip_fp16.h>
#include <stdint.h>

#include "lbm_d3q19.hpp"

namespace lbm {

// Policies whose encodings carry infinities and NaN (or that do not round)
// need no range check.
template <class Self>
struct Unchecked {
    static constexpr bool kChecked = false;
    __device__ static __forceinline__ bool overflowed(uint32_t) { return false; }
};

struct StoreF32 : Unchecked<StoreF32> {
    using S = float;
    using C = float;
    static constexpr bool kShifted = false;
    __device__ static __forceinline__ C load(S v, int) { return v; }
    __device__ static __forceinline__ S store(C v, int) { return v; }
    __device__ static __forceinline__ S store_checked(C v, int l, uint32_t&) { return store(v, l); }
};

struct StoreF64 : Unchecked<StoreF64> {
    using S = double;
    using C = double;
    static constexpr bool kShifted = false;
    __device__ static __forceinline__ C load(S v, int) { return v; }
    __device__ static __forceinline__ S store(C v, int) { return v; }
    __device__ static __forceinline__ S store_checked(C v, int l, uint32_t&) { return store(v, l); }
};

// FP16 keeps 11 significant bits only above 2^-14; slow regions of a cavity
// have |f - w| ~ 1e-5, which would fall into its subnormal range. The stored
// value is (f - w) * 2^12: exact scaling, normal range down to |f - w| ~ 1.5e-8,
// and |f - w| <= ~0.1 stays far below the FP16 maximum (65504).
struct StoreF16S : Unchecked<StoreF16S> {
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
    __device__ static __forceinline__ S store_checked(C v, int l, uint32_t&) { return store(v, l); }
};

struct StoreBF16S : Unchecked<StoreBF16S> {
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
    __device__ static __forceinline__ S store_checked(C v, int l, uint32_t&) { return store(v, l); }
};

// FP16C of Lehmann et al. (Phys. Rev. E 106, 015308, 2022): 1 sign, 4 exponent
// and 11 significand bits, here of x = f_l - w_l with the largest binade
// [2^-3, 2^-2). Code s|e|m: e = 1..15 encodes 2^(e - 18) (1 + m 2^-11),
// e = 0 the subnormal m 2^-28 (2^-28 is the FP32 step of f near w = 1/18, so
// FP32 arithmetic on w + x resolves no finer). One significand bit more than
// binary16 where shifted populations live (|x| stays below 0.03 in a cavity
// at lid speed 0.1); no infinities or NaN, so a store that rounds to
// |x| >= 2^-2, or of a non-finite value, is reported through store_checked.
//
// Encode: |y| = |x| * 2^-109 moves the exponent so that our e is the float's
// exponent field (2^(e - 18) * 2^-109 = 2^(e - 127)) and our subnormals are
// float subnormals; the multiplication is exact for every x = f - w with w >=
// 1/36 (x is then a multiple of 2^-29), and the 12 low bits of |y| are
// rounded to nearest even in integer arithmetic (the carry moves into e); the
// code is the sign of x over the rounded magnitude.
// Decode: the code shifted left by 12 is y again (sign-extending the 16-bit
// code copies the sign into bits 27..31, the mask keeps bit 31), and one FMA
// by 2^109 gives w + x with one rounding. FP32 subnormals are kept on gfx950
// (denormal mode "preserve"), which both directions rely on.
struct StoreF16C {
    using S = uint16_t;
    using C = float;
    static constexpr bool kShifted = true;
    static constexpr bool kChecked = true;
    static constexpr uint32_t kOverflow = 0x08000000u;   // rounded |y| >= 2^15 << 12
    __device__ static __forceinline__ C load(S v, int l)
    {
        const uint32_t y = ((uint32_t)(int32_t)(int16_t)v << 12) & 0x87FFFFFFu;
        return __builtin_fmaf(__uint_as_float(y), 0x1p109f, weight<float>(l));
    }
    __device__ static __forceinline__ S store_checked(C v, int l, uint32_t& ov)
    {
        const float x = v - weight<float>(l);
        const uint32_t a = __float_as_uint(__builtin_fabsf(x) * 0x1p-109f);
        const uint32_t r = a + 0x7FFu + ((a >> 12) & 1u);
        ov |= r;
        // sign in bit 31, the rounded magnitude in bits 16..30: the code is
        // the high half
        const uint32_t hi = (__float_as_uint(x) & 0x80000000u) | (r << 4);
        return (S)(hi >> 16);
    }
    __device__ static __forceinline__ S store(C v, int l)
    {
        uint32_t ov = 0;
        return store_checked(v, l, ov);
    }
    __device__ static __forceinline__ bool overflowed(uint32_t ov) { return ov >= kOverflow; }
};

// Marks the solver's state as out of the storage range (F16C only; a no-op for
// the other formats, whose encodings carry infinities and NaN). Every writer
// stores the same value, so concurrent raises need no atomics.
template <class P>
__device__ __forceinline__ void raise_overflow(uint32_t* status, uint32_t ov)
{
    if (P::kChecked && P::overflowed(ov) && status)
        *status = 1u;
}

}  // namespace lbm
