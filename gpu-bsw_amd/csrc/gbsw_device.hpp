// gbsw_device.hpp -- device-side building blocks of the gfx950 Smith-Waterman engine.
//
// Packed int16 arithmetic: every 32-bit VGPR holds two 16-bit DP cells that belong
// to two independent "virtual lanes" (see gbsw_kernels.hip), so each v_pk_* op
// advances two cells.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gbsw {

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_s(uint32_t v) { return __builtin_bit_cast(s16x2, v); }
__device__ __forceinline__ u16x2 as_u(uint32_t v) { return __builtin_bit_cast(u16x2, v); }
__device__ __forceinline__ uint32_t raw(s16x2 v) { return __builtin_bit_cast(uint32_t, v); }
__device__ __forceinline__ uint32_t raw(u16x2 v) { return __builtin_bit_cast(uint32_t, v); }

// v_pk_add_i16 clamp: signed saturating add of both halves.
__device__ __forceinline__ uint32_t pk_add_sat(uint32_t a, uint32_t b)
{
    return raw(__builtin_elementwise_add_sat(as_s(a), as_s(b)));
}
// v_pk_max_i16
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b)
{
    return raw(__builtin_elementwise_max(as_s(a), as_s(b)));
}
// v_pk_sub_u16 clamp: max(a - b, 0) per half for non-negative a, b.
__device__ __forceinline__ uint32_t pk_sub_floor0(uint32_t a, uint32_t b)
{
    return raw(__builtin_elementwise_sub_sat(as_u(a), as_u(b)));
}
// 0xFFFF in each half where a < b (signed; |a - b| < 2^15), else 0:
// v_pk_sub_i16 + v_pk_ashrrev_i16.  Written as asm: in the int16 kernels the
// compiler's own form of this sign test measured 3 % slower.
__device__ __forceinline__ uint32_t pk_lt_mask(uint32_t a, uint32_t b)
{
    uint32_t d, m;
    asm("v_pk_sub_i16 %0, %1, %2" : "=v"(d) : "v"(a), "v"(b));
    asm("v_pk_ashrrev_i16 %0, 15, %1 op_sel_hi:[0,1]" : "=v"(m) : "v"(d));
    return m;
}
// v_bfi_b32: (m & a) | (~m & b).  Written as asm for the same reason (the
// compiler's form of the snapshot selects measured 14 % slower overall).
__device__ __forceinline__ uint32_t bfi(uint32_t m, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
    return r;
}
// ---- integer-valued packed f16 (exact for |x| <= 2048)
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ f16x2 as_h(uint32_t v) { return __builtin_bit_cast(f16x2, v); }
__device__ __forceinline__ uint32_t raw(f16x2 v) { return __builtin_bit_cast(uint32_t, v); }
// v_pk_add_f16
__device__ __forceinline__ uint32_t pk_add_f16(uint32_t a, uint32_t b) { return raw(as_h(a) + as_h(b)); }
// IEEE maximum of both halves (llvm.maximum: v_pk_maximum3_f16 on gfx950, with no
// canonicalising max per operand, unlike maxnum).  The operands are integers,
// never NaN, and -0 does not arise (a sum is -0 only if both terms are).
__device__ __forceinline__ uint32_t pk_max_f16(uint32_t a, uint32_t b)
{
    return raw(__builtin_elementwise_maximum(as_h(a), as_h(b)));
}
// v_pk_maximum3_f16 (gfx950): max of three per half.
__device__ __forceinline__ uint32_t pk_max3_f16(uint32_t a, uint32_t b, uint32_t c)
{
    return raw(__builtin_elementwise_maximum(__builtin_elementwise_maximum(as_h(a), as_h(b)), as_h(c)));
}
// 0xFFFF in each half where a < b, else 0: the sign of a - b (exact, and +0 when
// equal), spread by v_pk_ashrrev_i16.
__device__ __forceinline__ uint32_t pk_lt_mask_f16(uint32_t a, uint32_t b)
{
    return raw(as_s(raw(as_h(a) - as_h(b))) >> (s16x2){15, 15});
}
__host__ __device__ inline uint16_t f16_bits(int32_t x) { return __builtin_bit_cast(uint16_t, (_Float16)x); }
__device__ __forceinline__ int32_t f16_int(uint32_t bits16)
{
    return (int32_t)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}

// v_bfi_b32 with a wave-uniform (SGPR) second operand.
__device__ __forceinline__ uint32_t bfi_s(uint32_t m, uint32_t a, uint32_t b)
{
    uint32_t r;
    asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "s"(a), "v"(b));
    return r;
}
// Rotate right by one lane inside each 16-lane row (DPP row_ror:1).  Every lane
// reads a valid source, so the destination needs no initial value.
__device__ __forceinline__ uint32_t row_ror1(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x121, 0xf, 0xf, false);
}
__device__ __forceinline__ int32_t lo16(uint32_t v) { return (int32_t)(int16_t)(v & 0xFFFFu); }
__device__ __forceinline__ int32_t hi16(uint32_t v) { return (int32_t)(int16_t)(v >> 16); }
__device__ __forceinline__ uint32_t pack16(int32_t lo, int32_t hi)
{
    return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16);
}

// Tie priority of column t in the reference's forward arg-max tree
// (kernel.cpp:47-86, 137-182): lanes compare in shfl_down order, so the survivor
// of a full tie has the smallest bit-reversed (warp, lane) index.  Columns past
// 1024 (no reference counterpart) order by t>>10 first.
__device__ __forceinline__ uint32_t tie_prio(uint32_t t)
{
    const uint32_t lane = __builtin_bitreverse32(t & 31u) >> 27;
    const uint32_t warp = __builtin_bitreverse32((t >> 5) & 31u) >> 27;
    return ((t >> 10) << 10) | (warp << 5) | lane;
}

} // namespace gbsw
