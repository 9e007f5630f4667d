// gbsw_kernels.hip -- gfx950 kernels of the batched Smith-Waterman engine.
//
// Replaces the reference's sequence_{dna,aa}_{kernel,reverse} (kernel.cpp:198-1059)
// with one wavefront engine for both alphabets.
//
// Geometry.  A wave64 aligns four pairs, one per 16-lane DPP row ("group").  Each
// lane owns two column blocks of K columns of the shorter sequence (the
// reference's column orientation, kernel.cpp:248-259): the low int16 half of every
// state register holds block l ("virtual lane" v = l), the high half block l+16
// (v = l+16).  Virtual lane v works on row s-v at step s, so the 32 virtual lanes
// of a pair form an anti-diagonal wavefront over a strip of W = 32*K columns, and
// the two halves of a register are always independent cells: every v_pk_* op
// advances two DP cells.  Longer column sequences are processed strip by strip;
// strip t+1 reads the last column of strip t (H and E per row) from global memory.
//
// Recurrence (same H as kernel.cpp:329-367; E and F are kept clamped at 0, which
// cannot change H = max(D, E, F, 0) when the gap scores are <= 0):
//   D  = H(i-1,j-1) + s(row, col)                      v_pk_add_i16 clamp
//   H  = max(D, E', F')                                2 x v_pk_max_i16
//   Ho = max(H + open, 0)                              v_pk_sub_u16 clamp
//   E' = max(max(E'+ext, 0), Ho)   (next column)       v_pk_sub_u16 clamp + max
//   F' = max(max(F'+ext, 0), Ho)   (next row)          v_pk_sub_u16 clamp + max
// Substitution scores come from a per-strip profile in LDS: row a of the profile
// holds s(code a, column j) for the strip's columns; row A is the padding row.
//
// Results follow the reference's tie rules exactly:
//  forward: among cells with H = score, the smallest row, then the smallest tree
//    priority of the column (kernel.cpp:47-86, 137-182, 368-370).  Each virtual lane
//    keeps (best, first row reaching it, snapshot of that row); the winner is
//    chosen after the pass.
//  reverse: the same DP on the reversed prefixes; among cells with H = score, the
//    smallest reversed column, then its first row (kernel.cpp:3-45, 577-581).
#include <type_traits>

#include "gbsw_device.hpp"
#include "gbsw_kernel_args.hpp"
#include "gbsw_kernels.hpp"

namespace gbsw {

// One pass (forward or reverse) of one pair, identical in the 16 lanes of a group.
struct Task {
    const uint8_t* rows;      // row sequence (the longer one of this pass)
    const uint8_t* cols;      // column sequence
    int32_t nr, nc;           // rows / columns computed
    int32_t nr_full, nc_full; // extents the reversed indices count back from
    int32_t rev;              // 1: both sequences are read backwards (reverse pass)
};

struct Smem {
    uint8_t* base;
    LdsLayout L;
    __device__ uint8_t* lut() const { return base + L.lut_off; }
    __device__ int16_t* sub() const { return (int16_t*)(base + L.sub_off); }
    __device__ int16_t* colmax() const { return (int16_t*)(base + L.colmax_off); }
    __device__ uint64_t* xch() const { return (uint64_t*)(base + L.xch_off); }
    __device__ int32_t* info() const { return (int32_t*)(base + L.info_off); }
    __device__ uint8_t* group(int g) const { return base + L.group_off + g * L.group_bytes; }
};

__device__ __forceinline__ uint32_t inv_prio(uint32_t p)
{
    // tie_prio is a per-1024-block bit-reversal of two 5-bit fields: its own inverse.
    return tie_prio(p);
}

// ---------------------------------------------------------------- arithmetic

// Per-lane constants of the lane-0 fix-up after the row rotation: in lane 0 the
// low virtual lane (0) takes the strip boundary and the high one (16) takes the
// rotated low half (virtual lane 15); other lanes keep the rotated value.
struct LaneFix {
    uint32_t mask; // all ones in lane 0
    uint32_t selh; // v_perm selector: lane 0 -> (boundary H, rotated low), else identity
    uint32_t sele; // the same with the boundary's E half
};
__device__ __forceinline__ LaneFix lane_fix(int l)
{
    LaneFix f;
    f.mask = l == 0 ? 0xFFFFFFFFu : 0u;
    f.selh = l == 0 ? 0x05040100u : 0x07060504u;
    f.sele = l == 0 ? 0x05040302u : 0x07060504u;
    return f;
}

// Packed16: one VGPR holds the cells of both virtual lanes of a lane as int16
// halves (low: virtual lane l, high: l+16); every op advances two cells.
struct Packed16 {
    using V = uint32_t;
    using B = uint32_t; // strip boundary entry: H | E << 16
    static constexpr bool kWide = false;
    __device__ static V zero() { return 0u; }
    __device__ static V splat(int32_t x) { return pack16(x, x); }
    // step numbers recorded in both halves (step < 2^16)
    __device__ static V sel_step(V m, int s, V b) { return bfi_s(m, pack16(s, s), b); }
    __device__ static int32_t step_lo(V r) { return (int32_t)(r & 0xFFFFu); }
    __device__ static int32_t step_hi(V r) { return (int32_t)(r >> 16); }
    __device__ static V dpp(V x) { return row_ror1(x); }
    // virtual lane 0 takes the strip boundary, virtual lane 16 the low half of lane 15
    __device__ static V fix_h(V x, B bnd, const LaneFix& f) { return __builtin_amdgcn_perm(x, bnd, f.selh); }
    __device__ static V fix_e(V x, B bnd, const LaneFix& f) { return __builtin_amdgcn_perm(x, bnd, f.sele); }
    __device__ static B bnd_out(V h, V e) { return __builtin_amdgcn_perm(e, h, 0x07060302u); } // high halves
    template <int K>
    __device__ static V prof(const uint32_t* plo, const uint32_t* phi, int k)
    {
        return (k & 1) ? __builtin_amdgcn_perm(phi[k / 2], plo[k / 2], 0x07060302u)
                       : __builtin_amdgcn_perm(phi[k / 2], plo[k / 2], 0x05040100u);
    }
    __device__ static V add(V a, V b) { return pk_add_sat(a, b); }
    __device__ static V vmax(V a, V b) { return pk_max(a, b); }
    __device__ static V vmax3(V a, V b, V c) { return pk_max(pk_max(a, b), c); }
    __device__ static V subf0(V a, V b) { return pk_sub_floor0(a, b); }
    // gap step: operand g(score) and the three uses of the recurrence
    __device__ static V gap(int32_t score) { return splat(-score); }
    __device__ static V cell(V D, V E, V F) { return pk_max(pk_max(D, E), F); }   // H
    __device__ static V open_in(V H, V G) { return pk_sub_floor0(H, G); }          // max(H + open, 0)
    __device__ static V decay(V X, V G, V Ho) { return pk_max(pk_sub_floor0(X, G), Ho); }
    __device__ static V ltmask(V a, V b) { return pk_lt_mask(a, b); }
    __device__ static V sel(V m, V a, V b) { return bfi(m, a, b); }
    __device__ static int32_t lo(V x) { return lo16(x); }
    __device__ static int32_t hi(V x) { return hi16(x); }
    // equality with a splat value, on the raw halves
    __device__ static bool eq_lo(V x, V t) { return (uint16_t)x == (uint16_t)t; }
    __device__ static bool eq_hi(V x, V t) { return (uint16_t)(x >> 16) == (uint16_t)(t >> 16); }
    // Some half of x reaches the target, given x <= target in both halves and t1
    // = the target's bits minus one in both halves: cell values are >= 0, where
    // the raw 16-bit patterns (int16 and f16 alike) order as the values do, so
    // one saturating unsigned subtraction and one compare decide it.
    __device__ static V target_less1(V t) { return t - 0x00010001u; }
    __device__ static bool reach_any(V x, V t1) { return raw(__builtin_elementwise_sub_sat(as_u(x), as_u(t1))) != 0u; }
    // some half differs (raw bits; cell values order as their bits)
    __device__ static bool differs(V a, V b) { return (a ^ b) != 0u; }
    // 1 in each half equal to t's (raw bits), else 0
    __device__ static uint32_t eq_bits(V x, V t)
    {
        const u16x2 one = {1, 1};
        return raw(one - __builtin_elementwise_min(as_u(x ^ t), one));
    }
    __device__ static int16_t prof_val(int32_t s) { return (int16_t)s; }
    // byte-profile entry of score s (kNeg, the padding, becomes -128)
    __device__ static uint8_t prof_byte(int32_t s) { return (uint8_t)(int8_t)max(s, -128); }
    // column k of a byte profile: byte k%4 of word k/4, sign-extended into each
    // half.  v_perm's selectors 8..11 copy the sign bits of bytes 1, 3, 5, 7, so
    // odd bytes take one instruction; even ones go to the high byte of each half
    // and shift down.
    template <int K>
    __device__ static V prof8(const uint32_t* plo, const uint32_t* phi, int k)
    {
        const uint32_t wl = plo[k / 4], wh = phi[k / 4];
        switch (k & 3) {
        case 1: return __builtin_amdgcn_perm(wh, wl, 0x0A050801u);
        case 3: return __builtin_amdgcn_perm(wh, wl, 0x0B070903u);
        case 0: return raw(as_s(__builtin_amdgcn_perm(wh, wl, 0x040C000Cu)) >> (s16x2){8, 8});
        default: return raw(as_s(__builtin_amdgcn_perm(wh, wl, 0x060C020Cu)) >> (s16x2){8, 8});
        }
    }
    __device__ static B load_bnd(const B* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
    __device__ static B zero_bnd() { return 0u; }
};

// PackedF16: the Packed16 layout with each half an integer-valued f16, exact
// while every value stays within +-2048.  The host uses it only for length
// buckets whose largest possible score (shorter length x best substitution)
// plus one substitution is below 2048; there H = max(D, E, F) and the max over a
// lane's columns take one v_pk_maximum3_f16 per three operands.  E and F keep
// their floor at 0 through the third operand, so every cell value equals the
// Packed16 one.  Very negative D values (padding, -16384 + H) are rounded but
// never selected, since E, F >= 0.
struct PackedF16 : Packed16 {
    __device__ static V splat(int32_t x)
    {
        const uint32_t b = f16_bits(x);
        return b | (b << 16);
    }
    __device__ static V add(V a, V b) { return pk_add_f16(a, b); }
    __device__ static V vmax(V a, V b) { return pk_max_f16(a, b); }
    __device__ static V vmax3(V a, V b, V c) { return pk_max3_f16(a, b, c); }
    __device__ static V gap(int32_t score) { return splat(score); }
    __device__ static V cell(V D, V E, V F) { return pk_max3_f16(D, E, F); }
    __device__ static V open_in(V H, V G) { return pk_add_f16(H, G); } // may be < 0: decay floors it
    __device__ static V decay(V X, V G, V Ho) { return pk_max3_f16(pk_add_f16(X, G), Ho, 0u); }
    __device__ static V ltmask(V a, V b) { return pk_lt_mask_f16(a, b); }
    __device__ static int32_t lo(V x) { return f16_int(x & 0xFFFFu); }
    __device__ static int32_t hi(V x) { return f16_int(x >> 16); }
    __device__ static int16_t prof_val(int32_t s) { return (int16_t)f16_bits(s); }
    // Byte profile of a table whose scores (and kNeg) are f16 values with a zero
    // low byte: the high byte, and column k of it is that byte moved into the
    // high byte of each half (one v_perm: selector 0x0C is a zero byte).
    __device__ static uint8_t prof_byte(int32_t s) { return (uint8_t)(f16_bits(s) >> 8); }
    template <int K>
    __device__ static V prof8(const uint32_t* plo, const uint32_t* phi, int k)
    {
        const uint32_t b = (uint32_t)(k & 3);
        return __builtin_amdgcn_perm(phi[k / 4], plo[k / 4], 0x000C000Cu | (b << 8) | ((4u + b) << 24));
    }
};

// Wide32: the same two virtual lanes as two int32 VGPRs, for pairs whose score
// may exceed the int16 range.
struct Wide32 {
    struct V {
        int32_t l, h;
    };
    using B = uint2; // (H, E)
    static constexpr bool kWide = true;
    __device__ static V zero() { return V{0, 0}; }
    __device__ static V splat(int32_t x) { return V{x, x}; }
    __device__ static V sel_step(V m, int s, V b) { return V{m.l ? s : b.l, m.h ? s : b.h}; }
    __device__ static int32_t step_lo(V r) { return r.l; }
    __device__ static int32_t step_hi(V r) { return r.h; }
    __device__ static V dpp(V x) { return V{(int32_t)row_ror1((uint32_t)x.l), (int32_t)row_ror1((uint32_t)x.h)}; }
    __device__ static V fix_h(V x, B bnd, const LaneFix& f)
    {
        return V{(int32_t)bfi(f.mask, bnd.x, (uint32_t)x.l), (int32_t)bfi(f.mask, (uint32_t)x.l, (uint32_t)x.h)};
    }
    __device__ static V fix_e(V x, B bnd, const LaneFix& f)
    {
        return V{(int32_t)bfi(f.mask, bnd.y, (uint32_t)x.l), (int32_t)bfi(f.mask, (uint32_t)x.l, (uint32_t)x.h)};
    }
    __device__ static B bnd_out(V h, V e) { return make_uint2((uint32_t)h.h, (uint32_t)e.h); }
    template <int K>
    __device__ static V prof(const uint32_t* plo, const uint32_t* phi, int k)
    {
        const int sh = (k & 1) * 16;
        return V{(int32_t)(int16_t)(plo[k / 2] >> sh), (int32_t)(int16_t)(phi[k / 2] >> sh)};
    }
    __device__ static V add(V a, V b) { return V{a.l + b.l, a.h + b.h}; }
    __device__ static V vmax(V a, V b) { return V{max(a.l, b.l), max(a.h, b.h)}; }
    __device__ static V vmax3(V a, V b, V c) { return vmax(vmax(a, b), c); }
    __device__ static V subf0(V a, V b)
    {
        return V{(int32_t)__builtin_elementwise_sub_sat((uint32_t)a.l, (uint32_t)b.l),
                 (int32_t)__builtin_elementwise_sub_sat((uint32_t)a.h, (uint32_t)b.h)};
    }
    __device__ static V gap(int32_t score) { return splat(-score); }
    __device__ static V cell(V D, V E, V F) { return vmax(vmax(D, E), F); }
    __device__ static V open_in(V H, V G) { return subf0(H, G); }
    __device__ static V decay(V X, V G, V Ho) { return vmax(subf0(X, G), Ho); }
    __device__ static bool eq_lo(V x, V t) { return x.l == t.l; }
    __device__ static bool eq_hi(V x, V t) { return x.h == t.h; }
    __device__ static V target_less1(V t) { return t; }
    __device__ static bool reach_any(V x, V t) { return x.l == t.l || x.h == t.h; }
    __device__ static bool differs(V a, V b) { return a.l != b.l || a.h != b.h; }
    __device__ static int16_t prof_val(int32_t s) { return (int16_t)s; }
    __device__ static uint8_t prof_byte(int32_t s) { return (uint8_t)(int8_t)max(s, -128); }
    template <int K>
    __device__ static V prof8(const uint32_t* plo, const uint32_t* phi, int k)
    {
        const int sh = (k & 3) * 8;
        return V{(int32_t)(int8_t)(plo[k / 4] >> sh), (int32_t)(int8_t)(phi[k / 4] >> sh)};
    }
    __device__ static V ltmask(V a, V b) { return V{a.l < b.l ? -1 : 0, a.h < b.h ? -1 : 0}; }
    __device__ static V sel(V m, V a, V b)
    {
        return V{(int32_t)bfi((uint32_t)m.l, (uint32_t)a.l, (uint32_t)b.l),
                 (int32_t)bfi((uint32_t)m.h, (uint32_t)a.h, (uint32_t)b.h)};
    }
    __device__ static int32_t lo(V x) { return x.l; }
    __device__ static int32_t hi(V x) { return x.h; }
    __device__ static B load_bnd(const B* p)
    {
        const uint64_t v = __hip_atomic_load((const uint64_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_uint2((uint32_t)v, (uint32_t)(v >> 32));
    }
    __device__ static B zero_bnd() { return make_uint2(0u, 0u); }
};

// Raw bytes of 2 ring rows prefetched 32 steps before they are stored in LDS.
template <class A>
struct RowFetch {
    int32_t b0, b1;         // residue byte, -1 outside [0, nr)
    typename A::B d0, d1;   // strip boundary (H, E)
};

template <class A>
__device__ __forceinline__ RowFetch<A> row_fetch(const Task& t, int chunk, int l, const typename A::B* bnd, bool use_bnd)
{
    RowFetch<A> f;
    const int x0 = chunk * kChunk + l, x1 = x0 + 16;
    const bool v0 = x0 >= 0 && x0 < t.nr, v1 = x1 >= 0 && x1 < t.nr;
    f.b0 = v0 ? (int32_t)t.rows[t.rev ? t.nr_full - 1 - x0 : x0] : -1;
    f.b1 = v1 ? (int32_t)t.rows[t.rev ? t.nr_full - 1 - x1 : x1] : -1;
    // the boundary rows were stored by this workgroup during the previous strip:
    // an agent-scope load does not hit a stale L1 line
    f.d0 = (v0 && use_bnd) ? A::load_bnd(bnd + x0) : A::zero_bnd();
    f.d1 = (v1 && use_bnd) ? A::load_bnd(bnd + x1) : A::zero_bnd();
    return f;
}

template <class A>
__device__ __forceinline__ void row_store(const RowFetch<A>& f, int chunk, int l, uint16_t* ring_code,
                                          typename A::B* ring_bnd, const uint8_t* lut, int alpha, int rowstride)
{
    const int x0 = chunk * kChunk + l, x1 = x0 + 16;
    int c0 = f.b0 < 0 ? alpha : (int)lut[f.b0];
    int c1 = f.b1 < 0 ? alpha : (int)lut[f.b1];
    c0 = c0 > alpha ? alpha : c0; // bytes outside the alphabet read the padding row
    c1 = c1 > alpha ? alpha : c1;
    const int i0 = x0 & (kRing - 1), i1 = x1 & (kRing - 1);
    ring_code[i0] = (uint16_t)(c0 * rowstride);
    ring_code[i1] = (uint16_t)(c1 * rowstride);
    // the copy after the ring's end
    if (i0 < kRingTail) ring_code[i0 + kRing] = (uint16_t)(c0 * rowstride);
    if (i1 < kRingTail) ring_code[i1 + kRing] = (uint16_t)(c1 * rowstride);
    ring_bnd[x0 & (kRing - 1)] = f.d0;
    ring_bnd[x1 & (kRing - 1)] = f.d1;
}

// Load the kstride(K) profile scores of one virtual lane v from profile row
// `row`: K/8 b128 blocks (block q at q*512 + v*16 bytes) and the remainder at
// (K/8)*512 + v*kremstride(K)*2.  Dword d holds columns 2d and 2d+1.
template <int K>
__device__ __forceinline__ void prof_load(const uint8_t* row, int v, uint32_t (&d)[kstride(K) / 2])
{
    constexpr int Q = K / 8, RS = kremstride(K);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint4 x = *(const uint4*)(row + q * 512 + v * 16);
        d[4 * q + 0] = x.x; d[4 * q + 1] = x.y; d[4 * q + 2] = x.z; d[4 * q + 3] = x.w;
    }
    const uint8_t* r = row + Q * 512 + v * RS * 2;
    if constexpr (RS == 8) {
        const uint4 x = *(const uint4*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y; d[4 * Q + 2] = x.z; d[4 * Q + 3] = x.w;
    } else if constexpr (RS == 6) {
        const uint3 x = *(const uint3*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y; d[4 * Q + 2] = x.z;
    } else if constexpr (RS == 4) {
        const uint2 x = *(const uint2*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y;
    } else if constexpr (RS == 2) {
        d[4 * Q] = *(const uint32_t*)r;
    }
}

// Byte profile: the kstride8(K) bytes of virtual lane v, as kstride8(K)/4 dwords
// (dword d holds columns 4d .. 4d+3).
template <int K>
__device__ __forceinline__ void prof_load8(const uint8_t* row, int v, uint32_t (&d)[kstride8(K) / 4])
{
    constexpr int Q = K / 16, RS = kremstride8(K);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint4 x = *(const uint4*)(row + q * 512 + v * 16);
        d[4 * q + 0] = x.x; d[4 * q + 1] = x.y; d[4 * q + 2] = x.z; d[4 * q + 3] = x.w;
    }
    const uint8_t* r = row + Q * 512 + v * RS;
    if constexpr (RS == 16) {
        const uint4 x = *(const uint4*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y; d[4 * Q + 2] = x.z; d[4 * Q + 3] = x.w;
    } else if constexpr (RS == 12) {
        const uint3 x = *(const uint3*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y; d[4 * Q + 2] = x.z;
    } else if constexpr (RS == 8) {
        const uint2 x = *(const uint2*)r;
        d[4 * Q + 0] = x.x; d[4 * Q + 1] = x.y;
    } else if constexpr (RS == 4) {
        d[4 * Q] = *(const uint32_t*)r;
    }
}

// Byte-profile offset j (0..32*kstride8(K)) -> (virtual lane, column within it).
template <int K>
__device__ __forceinline__ void prof_slot8(int j, int& v, int& k)
{
    constexpr int Q = K / 16, RS = kremstride8(K);
    if (j < Q * 512) {
        v = (j & 511) >> 4;
        k = (j >> 9) * 16 + (j & 15);
    } else {
        const int jj = j - Q * 512;
        v = jj / RS;
        k = Q * 16 + (jj - v * RS);
    }
}

// Profile column index j (0..32*kstride(K)) -> (virtual lane, column within it);
// k >= K marks padding.
template <int K>
__device__ __forceinline__ void prof_slot(int j, int& v, int& k)
{
    constexpr int Q = K / 8, RS = kremstride(K);
    if (j < Q * 256) {
        const int q = j >> 8;
        v = (j & 255) >> 3;
        k = q * 8 + (j & 7);
    } else {
        const int jj = j - Q * 256;
        v = RS ? jj / RS : 0;
        k = Q * 8 + (RS ? jj - v * RS : 0);
    }
}

// Max over K registers with ceil((K - 1) / 2) three-input maxima: a queue, each
// maximum takes the three oldest entries and appends its result (depth about
// log3 K).  Op j reads entries 3j .. 3j + 2 of the queue, which then holds
// K - 2j entries; only a last op with two entries left is a two-input maximum.
template <class A, int K>
__device__ __forceinline__ typename A::V tree_max(const typename A::V (&h)[K])
{
    constexpr int NOPS = (K - 1 + 1) / 2;
    typename A::V q[K + NOPS];
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = h[k];
#pragma unroll
    for (int j = 0; j < NOPS; ++j)
        q[K + j] = (K - 2 * j >= 3) ? A::vmax3(q[3 * j], q[3 * j + 1], q[3 * j + 2]) : A::vmax(q[3 * j], q[3 * j + 1]);
    return q[K + NOPS - 1];
}

// Bits 0 .. n-1 set (n clamped to [0, 16]): the columns of a block inside the pair.
__device__ __forceinline__ uint32_t cols_mask(int n)
{
    return n <= 0 ? 0u : n >= 16 ? 0xFFFFu : (1u << n) - 1u;
}

struct PassResult {
    uint64_t key; // 0: no cell reached a positive score (forward) / the target (reverse)
};

// Runs one pass for the four groups of the wave.  All loops are wave-uniform;
// a group whose pair is shorter keeps computing on padding rows and columns,
// which can neither exceed nor (by row/column order) win over its real cells.
//
// Forward key: (score << 40) | ((2^20-1-row) << 20) | (2^20-1-prio(col)), max wins.
// Reverse key: ((2^20-1-col) << 20) | (2^20-1-row), max wins (target = score).
// One strip of the split form: its boundary input (the previous strip's rows,
// with that slot's progress word) and output.
struct StripIO {
    int strip;
    const void* bnd_in;
    void* bnd_out;
    const uint64_t* prog_in;
    uint64_t* prog_out;
    uint64_t epoch_in, epoch_out; // epoch << 32
};

// Wave-uniform wait until *w >= target.  Bounded: after kSpinTicks the timeout
// status word is set and the wave goes on (the host then reports the batch as
// failed), and once that word is set every wait returns at once, so a stalled
// pipeline drains instead of timing out wait by wait.
template <class T>
__device__ __forceinline__ void wait_at_least(const T* w, T target, int32_t* status)
{
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    for (;;) {
        uint64_t v = 0;
        int32_t stop = 0;
        if (threadIdx.x == 0) {
            v = (uint64_t)__hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            stop = __hip_atomic_load(status + kStatusTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
        if ((((uint64_t)hi << 32) | lo) >= (uint64_t)target) break;
        if (__builtin_amdgcn_readfirstlane(stop) != 0) break;
        if (__builtin_amdgcn_s_memrealtime() - t0 > kSpinTicks) {
            if (threadIdx.x == 0) atomicOr(status + kStatusTimeout, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}
// Per 16-lane group: wait until the group's word *w >= target (no wait where w
// is null) and return it (target where w is null).  Bounded like wait_at_least.
__device__ __forceinline__ uint32_t wait_group_u32(const uint32_t* w, uint32_t target, int32_t* status)
{
    const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
    uint32_t v = target;
    for (;;) {
        if (w) v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__ballot(v < target) == 0) break;
        int32_t stop = 0;
        if (threadIdx.x == 0) stop = __hip_atomic_load(status + kStatusTimeout, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (__builtin_amdgcn_readfirstlane(stop) != 0) break;
        if (__builtin_amdgcn_s_memrealtime() - t0 > kSpinTicks) {
            if (threadIdx.x == 0) atomicOr(status + kStatusTimeout, 1);
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    return v;
}
__device__ __forceinline__ void wait_word(const uint64_t* w, uint64_t target, int32_t* status)
{
    wait_at_least<uint64_t>(w, target, status);
}
__device__ __forceinline__ void wait_u32(const uint32_t* w, uint32_t target, int32_t* status)
{
    wait_at_least<uint32_t>(w, target, status);
}

// The wave's earlier global stores and atomics have completed (reached L2),
// then *w = v.  Readers on the same XCD see the stores once they see v.
__device__ __forceinline__ void publish_word(uint64_t* w, uint64_t v)
{
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (threadIdx.x == 0) __hip_atomic_store(w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// P8: the strip's profile holds int8 scores (prof_row_bytes(K, true) per row).
// SPLIT: run strip io.strip only, reading the previous strip's boundary rows from
// io.bnd_in as its producer publishes them and handing this strip's rows on
// through io.bnd_out; otherwise every strip, with the boundary in bnd_g.
template <class A, int K, bool REV, bool P8, bool SPLIT = false>
__device__ PassResult run_pass(const KernelArgs& a, const Smem& sm, const Task& t, int32_t target,
                               typename A::B* bnd_g, const StripIO& io = StripIO{})
{
    using V = typename A::V;
    using Bd = typename A::B;
    constexpr int KS = kstride(K);
    constexpr int NW = P8 ? kstride8(K) / 4 : KS / 2; // profile dwords of one virtual lane and row
    constexpr int W = kVirtPerPair * K;   // columns of one strip
    constexpr int WA = kVirtPerPair * KS; // profile columns reserved per row (int16 profile)
    const int lane = threadIdx.x;
    const int g = lane >> 4, l = lane & 15;
    const int alpha = a.alpha;
    constexpr int rowstride = prof_row_bytes(K, P8); // bytes between profile rows
    const LaneFix lfix = lane_fix(l);
    uint8_t* gb = sm.group(g);
    int16_t* prof = (int16_t*)gb;
    uint16_t* ring_code = (uint16_t*)(gb + sm.L.ring_code_off);
    Bd* ring_bnd = (Bd*)(gb + sm.L.ring_bnd_off);
    const uint8_t* lut = sm.lut();
    const int16_t* sub = sm.sub();

    // wave-uniform extents
    const int strips_g = (t.nr > 0 && t.nc > 0) ? (t.nc + W - 1) / W : 0;
    const int steps_g = t.nr + kVirtPerPair - 1;
    int32_t* info = sm.info();
    __syncthreads();
    if (l == 0) { info[g * 8 + 0] = strips_g; info[g * 8 + 1] = steps_g; }
    __syncthreads();
    int strips_max = 0, steps_max = 0;
#pragma unroll
    for (int q = 0; q < kPairsPerWave; ++q) {
        strips_max = max(strips_max, info[q * 8 + 0]);
        steps_max = max(steps_max, info[q * 8 + 1]);
    }
    // the same in every lane: keep the loop bounds scalar
    strips_max = __builtin_amdgcn_readfirstlane(strips_max);
    steps_max = __builtin_amdgcn_readfirstlane(steps_max);

    const bool multi_strip = strips_max > 1;
    const V OPEN2 = A::gap(a.gap_open);
    const V EXT2 = A::gap(a.gap_extend);
    const V TGT = A::splat(target); // reverse pass: the forward score (the box's maximum)
    const V TGT1 = A::target_less1(TGT);
    const int v_lo = l, v_hi = l + 16;
    uint64_t cand_lo = 0, cand_hi = 0;

    Bd* bnd_rd = SPLIT ? (Bd*)io.bnd_in : bnd_g;
    Bd* bnd_wr = SPLIT ? (Bd*)io.bnd_out : bnd_g;
    const int strip_lo = SPLIT ? io.strip : 0;
    const int strip_hi = SPLIT ? min(io.strip + 1, strips_max) : strips_max;
    for (int strip = strip_lo; strip < strip_hi; ++strip) {
        const int strip_base = strip * W;
        const bool live = strip < strips_g;
        const bool use_bnd = strip > 0;
        const bool write_bnd = live && strip + 1 < strips_g;
        // split form: boundary rows [0, n) of the previous strip are readable once
        // its progress word reaches epoch_in | n
        auto need_rows = [&](const int n) {
            if constexpr (SPLIT) {
                if (strip > 0) wait_word(io.prog_in, io.epoch_in | (uint64_t)(uint32_t)n, a.status);
            }
        };

        // ---- profile of this strip: row a holds s(code a, column) for the strip's
        // columns, the padding score for columns outside the pair and in row alpha
        __syncthreads();
        if constexpr (!P8) {
            for (int j = l; j < WA; j += 16) {
                int vj, kj;
                prof_slot<K>(j, vj, kj);
                const int c = strip_base + vj * K + kj;
                int code = -1;
                if (live && kj < K && c < t.nc) {
                    code = lut[t.cols[t.rev ? t.nc_full - 1 - c : c]];
                    if (code >= alpha) code = -1;
                }
                for (int r = 0; r <= alpha; ++r)
                    prof[r * WA + j] = A::prof_val((code >= 0 && r < alpha) ? sub[r * alpha + code] : kNeg);
            }
        } else {
            // four byte columns (one dword) per lane and row
            uint8_t* prof8 = (uint8_t*)gb;
            for (int j4 = l; j4 < rowstride / 4; j4 += 16) {
                int code[4];
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    int vj, kj;
                    prof_slot8<K>(j4 * 4 + b, vj, kj);
                    const int c = strip_base + vj * K + kj;
                    code[b] = -1;
                    if (live && kj < K && c < t.nc) {
                        code[b] = lut[t.cols[t.rev ? t.nc_full - 1 - c : c]];
                        if (code[b] >= alpha) code[b] = -1;
                    }
                }
                for (int r = 0; r <= alpha; ++r) {
                    uint32_t w = 0;
#pragma unroll
                    for (int b = 0; b < 4; ++b) {
                        const int sc = (code[b] >= 0 && r < alpha) ? sub[r * alpha + code[b]] : kNeg;
                        w |= (uint32_t)A::prof_byte(sc) << (8 * b);
                    }
                    *(uint32_t*)(prof8 + r * rowstride + j4 * 4) = w;
                }
            }
        }
        // ---- row ring: rows [-32, 0) padding, [0, 32) now, [32, 64) in flight
        {
            RowFetch<A> pad;
            pad.b0 = pad.b1 = -1;
            pad.d0 = pad.d1 = A::zero_bnd();
            row_store<A>(pad, -1, l, ring_code, ring_bnd, lut, alpha, rowstride);
            need_rows(2 * kChunk);
            RowFetch<A> f0 = row_fetch<A>(t, 0, l, bnd_rd, use_bnd);
            row_store<A>(f0, 0, l, ring_code, ring_bnd, lut, alpha, rowstride);
        }
        RowFetch<A> pf = row_fetch<A>(t, 1, l, bnd_rd, use_bnd);
        __syncthreads();


        V Hcol[K], Fcol[K];
        V snap[REV ? 1 : K];
#pragma unroll
        for (int k = 0; k < K; ++k) { Hcol[k] = A::zero(); Fcol[k] = A::zero(); }
#pragma unroll
        for (int k = 0; k < (REV ? 1 : K); ++k) snap[k] = A::zero();
        V Hout = A::zero(), Eout = A::zero(), Hdiag0 = A::zero();
        V best = A::zero(), bstep = A::zero(); // bstep: step of the first row reaching best
        int32_t rt_lo = 0x7FFFFFFF, rr_lo = 0, rt_hi = 0x7FFFFFFF, rr_hi = 0; // reverse hits

        // Rows [r0, r0 + 32) of the outgoing boundary, parked in their ring slots,
        // go to global memory: two coalesced stores per lane instead of one store
        // per step, issued after the next chunk's loads so that waiting for those
        // does not wait for the stores.  A slot is flushed (chunk after it was
        // written) before the row 128 further on enters it (two chunks later).
        auto flush_bnd = [&](const int r0) {
            if (write_bnd) {
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const int row = r0 + l + 16 * h;
                    if (row >= 0 && row < t.nr) bnd_wr[row] = ring_bnd[row & (kRing - 1)];
                }
            }
        };

        // waves of single-strip pairs read no boundary column and hand none on; the
        // two forms are separate loops so that neither holds the other's registers
        auto chunks = [&](auto ms) {
            for (int s0 = 0; s0 < steps_max; s0 += kChunk) {
                // refill: rows of chunk (s0/32 + 1) enter the ring, chunk +2 is requested
                {
                    const int ck = s0 / kChunk + 1;
                    row_store<A>(pf, ck, l, ring_code, ring_bnd, lut, alpha, rowstride);
                    need_rows((ck + 2) * kChunk);
                    pf = row_fetch<A>(t, ck + 1, l, bnd_rd, use_bnd);
                }
                __syncthreads();
                // rows leaving the last virtual lane during the previous chunk
                if (multi_strip && s0 > 0) {
                    flush_bnd(s0 - kChunk - (kVirtPerPair - 1));
                    if constexpr (SPLIT) {
                        // rows below s0 - 31 are stored once the flush completes: the
                        // wait for it (publish_word) is hidden by the other waves of
                        // the SIMD, and the next strip starts a chunk earlier than if
                        // the rows were published one chunk later
                        if (strip + 1 < strips_max)
                            publish_word(io.prog_out, io.epoch_out | (uint64_t)(uint32_t)(s0 - (kVirtPerPair - 1)));
                    }
                }
                // two steps per iteration (manual unroll: the DPP moves are convergent,
                // so the compiler will not unroll a runtime trip count); a step past
                // steps_max only meets padding rows
                const int s_end = min(s0 + kChunk, steps_max);
                // profile words and strip boundary of one step, requested a step ahead
                struct StepIn {
                    uint32_t plo[NW], phi[NW];
                    Bd bnd;
                };
                auto fetch = [&](StepIn& in, const uint32_t off_lo, const uint32_t off_hi, const Bd* pb, auto ms) {
                    if constexpr (P8) {
                        prof_load8<K>((const uint8_t*)prof + off_lo, v_lo, in.plo);
                        prof_load8<K>((const uint8_t*)prof + off_hi, v_hi, in.phi);
                    } else {
                        prof_load<K>((const uint8_t*)prof + off_lo, v_lo, in.plo);
                        prof_load<K>((const uint8_t*)prof + off_hi, v_hi, in.phi);
                    }
                    // the strip's left boundary column (zero in the first strip)
                    if constexpr (decltype(ms)::value) in.bnd = *pb;
                    else in.bnd = A::zero_bnd();
                };
                auto step = [&](const int s, const StepIn& in, auto ms) {
                    const uint32_t* plo = in.plo;
                    const uint32_t* phi = in.phi;
                    const Bd bnd = in.bnd;
                    // left inputs: neighbour's outputs of the previous step; virtual lane 0
                    // takes the strip's left boundary, virtual lane 16 takes lane 15's low half
                    V xH = A::dpp(Hout), xE = A::dpp(Eout);
                    xH = A::fix_h(xH, bnd, lfix);
                    xE = A::fix_e(xE, bnd, lfix);

                    V E = xE, Hd = Hdiag0;
    #pragma unroll
                    for (int k = 0; k < K; ++k) {
                        V S;
                        if constexpr (P8) S = A::template prof8<K>(plo, phi, k);
                        else S = A::template prof<K>(plo, phi, k);
                        const V D = A::add(Hd, S);
                        Hd = Hcol[k];
                        const V H = A::cell(D, E, Fcol[k]);
                        const V Ho = A::open_in(H, OPEN2);
                        E = A::decay(E, EXT2, Ho);
                        Fcol[k] = A::decay(Fcol[k], EXT2, Ho);
                        Hcol[k] = H;
                    }
                    Hout = Hcol[K - 1];
                    Eout = E;
                    Hdiag0 = xH;

                    // the last virtual lane hands its column to the next strip through
                    // its row's ring slot (whose incoming entry was read 31 steps ago);
                    // flush_bnd writes them out a chunk at a time
                    if constexpr (decltype(ms)::value) {
                        if (write_bnd && l == 15) ring_bnd[(s - v_hi) & (kRing - 1)] = A::bnd_out(Hout, Eout);
                    }

                    if constexpr (!REV) {
                        // the new best over the lane's columns and its old best
                        // (one maximum tree); the snapshot of the step's columns is
                        // taken in the steps where some lane of the wave improves
                        // (about two in three on the DNA mix), skipped in the others
                        V q[K + 1];
    #pragma unroll
                        for (int k = 0; k < K; ++k) q[k] = Hcol[k];
                        q[K] = best;
                        const V nb = tree_max<A, K + 1>(q);
                        if (__ballot(A::differs(nb, best)) != 0) {
                            const V m = A::ltmask(best, nb);
                            bstep = A::sel_step(m, s, bstep);
    #pragma unroll
                            for (int k = 0; k < K; ++k) snap[k] = A::sel(m, Hcol[k], snap[k]);
                        }
                        best = nb;
                    } else {
                        const V rm = tree_max<A, K>(Hcol);
                        if (A::reach_any(rm, TGT1)) {
                            const int row_lo = s - v_lo, row_hi = s - v_hi;
                            // first (smallest) column of this row that reaches the target
                            int k_lo = K, k_hi = K;
                            if constexpr (!A::kWide && K <= 16) {
                                // bit k of each half: column k equals the target
                                uint32_t acc = 0;
    #pragma unroll
                                for (int k = K - 1; k >= 0; --k)
                                    acc = raw(as_u(acc) * (u16x2){2, 2} + as_u(A::eq_bits(Hcol[k], TGT)));
                                const uint32_t ml = (acc & 0xFFFFu) & cols_mask(t.nc - (strip_base + v_lo * K));
                                const uint32_t mh = (acc >> 16) & cols_mask(t.nc - (strip_base + v_hi * K));
                                k_lo = ml ? __builtin_ctz(ml) : K;
                                k_hi = mh ? __builtin_ctz(mh) : K;
                            } else {
    #pragma unroll
                                for (int k = K - 1; k >= 0; --k) {
                                    if (A::eq_lo(Hcol[k], TGT) && strip_base + v_lo * K + k < t.nc) k_lo = k;
                                    if (A::eq_hi(Hcol[k], TGT) && strip_base + v_hi * K + k < t.nc) k_hi = k;
                                }
                            }
                            if (k_lo < K && row_lo >= 0 && row_lo < t.nr && strip_base + v_lo * K + k_lo < rt_lo) {
                                rt_lo = strip_base + v_lo * K + k_lo;
                                rr_lo = row_lo;
                            }
                            if (k_hi < K && row_hi >= 0 && row_hi < t.nr && strip_base + v_hi * K + k_hi < rt_hi) {
                                rt_hi = strip_base + v_hi * K + k_hi;
                                rr_hi = row_hi;
                            }
                        }
                    }
                };
                // this chunk's row codes are contiguous in the ring, those of virtual
                // lane v_lo 16 entries on from those of v_hi, and so are its boundary
                // entries (software pipeline: the LDS reads of step s + 1 are in flight
                // while step s computes)
                const uint16_t* rc0 = ring_code + ((s0 - v_hi) & (kRing - 1));
                const Bd* pb0 = ring_bnd + (s0 & (kRing - 1));
                auto steps = [&](auto ms) {
                    const uint16_t* rc = rc0;
                    const Bd* pb = pb0;
                    StepIn ia, ib;
                    fetch(ia, rc[16], rc[0], pb, ms);
                    for (int s = s0; s < s_end; s += 2) {
                        fetch(ib, rc[17], rc[1], pb + 1, ms);
                        step(s, ia, ms);
                        fetch(ia, rc[18], rc[2], pb + 2, ms);
                        step(s + 1, ib, ms);
                        rc += 2;
                        pb += 2;
                    }
                };
                steps(ms);
            }
        };
        if (multi_strip) chunks(std::true_type{});
        else chunks(std::false_type{});
        if (multi_strip) {
            // the last chunk's rows; complete before the next strip reads them
            __syncthreads();
            flush_bnd(((steps_max - 1) & ~(kChunk - 1)) - (kVirtPerPair - 1));
            if constexpr (SPLIT) {
                if (strip + 1 < strips_max) publish_word(io.prog_out, io.epoch_out | 0xFFFFFFFFull);
            } else {
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
            }
        }

        // ---- candidates of this strip
        if constexpr (!REV) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int32_t b = h ? A::hi(best) : A::lo(best);
                if (live && b > 0) {
                    const int32_t row = h ? A::step_hi(bstep) - v_hi : A::step_lo(bstep) - v_lo;
                    const int base = strip_base + (h ? v_hi : v_lo) * K;
                    uint32_t pbest = 0xFFFFFFFFu;
#pragma unroll
                    for (int k = 0; k < K; ++k) {
                        const int32_t v = h ? A::hi(snap[k]) : A::lo(snap[k]);
                        const int col = base + k;
                        if (v == b && col < t.nc) pbest = min(pbest, tie_prio((uint32_t)col));
                    }
                    if (pbest != 0xFFFFFFFFu && row < t.nr) {
                        const uint64_t key = ((uint64_t)b << 40) | ((uint64_t)(0xFFFFFu - (uint32_t)row) << 20) |
                                             (uint64_t)(0xFFFFFu - pbest);
                        if (h) cand_hi = key > cand_hi ? key : cand_hi;
                        else cand_lo = key > cand_lo ? key : cand_lo;
                    }
                }
            }
        } else {
            if (live && rt_lo != 0x7FFFFFFF) {
                const uint64_t key = ((uint64_t)(0xFFFFFu - (uint32_t)rt_lo) << 20) | (uint64_t)(0xFFFFFu - (uint32_t)rr_lo);
                cand_lo = key > cand_lo ? key : cand_lo;
            }
            if (live && rt_hi != 0x7FFFFFFF) {
                const uint64_t key = ((uint64_t)(0xFFFFFu - (uint32_t)rt_hi) << 20) | (uint64_t)(0xFFFFFu - (uint32_t)rr_hi);
                cand_hi = key > cand_hi ? key : cand_hi;
            }
            // a later strip only holds larger columns: stop once every group has a hit
            const uint64_t found = __ballot((cand_lo | cand_hi) != 0);
            bool all_done = true;
#pragma unroll
            for (int q = 0; q < kPairsPerWave; ++q) {
                const bool has = ((found >> (16 * q)) & 0xFFFFull) != 0;
                const bool more = info[q * 8 + 0] > strip + 1;
                all_done = all_done && (has || !more);
            }
            if (all_done) break;
        }
    }

    // ---- reduce the 32 candidates of each group
    uint64_t* xch = sm.xch();
    __syncthreads();
    xch[2 * lane + 0] = cand_lo;
    xch[2 * lane + 1] = cand_hi;
    __syncthreads();
    uint64_t key = 0;
    for (int q = 0; q < 2 * kLanesPerPair; ++q) {
        const uint64_t v = xch[g * 2 * kLanesPerPair + q];
        key = v > key ? v : key;
    }
    PassResult r;
    r.key = key;
    return r;
}

struct FwdOut {
    int32_t score, ref_end, query_end;
};

__device__ __forceinline__ Task empty_task(const KernelArgs& a, int rev)
{
    Task t;
    t.rows = a.ref; t.cols = a.qry;
    t.nr = t.nc = t.nr_full = t.nc_full = 0;
    t.rev = rev;
    return t;
}

__device__ __forceinline__ Task forward_task(const KernelArgs& a, int64_t p)
{
    if (p < 0) return empty_task(a, 0);
    Task t;
    t.rev = 0;
    const uint8_t* A = a.ref + a.ref_off[p];
    const uint8_t* B = a.qry + a.qry_off[p];
    const int32_t la = (int32_t)(a.ref_off[p + 1] - a.ref_off[p]);
    const int32_t lb = (int32_t)(a.qry_off[p + 1] - a.qry_off[p]);
    if (la < lb) { t.rows = B; t.nr = lb; t.cols = A; t.nc = la; } // kernel.cpp:248-253
    else { t.rows = A; t.nr = la; t.cols = B; t.nc = lb; }
    t.nr_full = t.nr;
    t.nc_full = t.nc;
    return t;
}

// Decodes the forward key.  A score at or above sat_limit may have saturated the
// int16 cells: the pair is queued for the int32 kernel (when record_redo) and
// reported as score -1.
__device__ __forceinline__ FwdOut forward_result(const KernelArgs& a, int64_t p, uint64_t key, bool record_redo = true)
{
    FwdOut o{0, 0, 0};
    if (p < 0 || key == 0) return o;
    const int32_t la = (int32_t)(a.ref_off[p + 1] - a.ref_off[p]);
    const int32_t lb = (int32_t)(a.qry_off[p + 1] - a.qry_off[p]);
    const int32_t score = (int32_t)(key >> 40);
    const int32_t row = (int32_t)(0xFFFFFu - (uint32_t)((key >> 20) & 0xFFFFFu));
    const int32_t col = (int32_t)inv_prio(0xFFFFFu - (uint32_t)(key & 0xFFFFFu));
    if (score >= a.sat_limit) {
        if (record_redo && (threadIdx.x & 15) == 0) {
            const int32_t slot = atomicAdd(a.status + kStatusSaturated, 1);
            a.redo[slot] = (int32_t)p;
        }
        o.score = -1;
        return o;
    }
    o.score = score;
    if (la < lb) { o.query_end = row + 1; o.ref_end = col + 1; } // kernel.cpp:381-386
    else { o.ref_end = row + 1; o.query_end = col + 1; }        // kernel.cpp:387-392
    return o;
}

// Sum over the forward column prefix of each residue's best possible score: an
// upper bound of the substitution part of any alignment inside the prefix.
__device__ __forceinline__ int32_t column_bound(const KernelArgs& a, const Smem& sm, int64_t p, const FwdOut& f)
{
    const int l = threadIdx.x & 15;
    int32_t u = 0;
    if (p >= 0 && f.score > 0) {
        const int32_t la = (int32_t)(a.ref_off[p + 1] - a.ref_off[p]);
        const int32_t lb = (int32_t)(a.qry_off[p + 1] - a.qry_off[p]);
        const bool ref_rows = !(la < lb);
        const uint8_t* C = ref_rows ? a.qry + a.qry_off[p] : a.ref + a.ref_off[p];
        const int32_t ec = ref_rows ? f.query_end : f.ref_end;
        const uint8_t* lut = sm.lut();
        const int16_t* cm = sm.colmax();
        for (int32_t j = l; j < ec; j += 16) {
            const int code = lut[C[j]];
            u += code < a.alpha ? (int32_t)cm[code] : 0;
        }
    }
#pragma unroll
    for (int o = 8; o > 0; o >>= 1) u += __shfl_xor(u, o, 16);
    return u;
}

// Reverse pass on the prefixes A[0:ref_end) x B[0:query_end), read backwards;
// orientation by the two end coordinates as in kernel.cpp:444-459.
//
// Exact extent bound: the forward winner has the smallest row R* among all cells
// scoring S, so every score-S alignment inside the prefixes ends on R*, i.e. at
// index 0 of the reversed forward-row sequence.  Such an alignment uses at most
// ec columns on the diagonal and its vertical gaps cost at least gap_min per
// position, so it spans at most ec + (U - S) / gap_min forward-row positions,
// with U = column_bound.  Cells farther along that axis cannot reach S, and
// cells nearer do not depend on them, so the pass stops there.
__device__ __forceinline__ Task reverse_task(const KernelArgs& a, int64_t p, const FwdOut& f, int32_t ubound)
{
    if (p < 0 || f.score <= 0) return empty_task(a, 1);
    Task t;
    t.rev = 1;
    const uint8_t* A = a.ref + a.ref_off[p];
    const uint8_t* B = a.qry + a.qry_off[p];
    const bool rev_rows_a = !(f.ref_end < f.query_end);
    if (!rev_rows_a) { t.rows = B; t.nr = f.query_end; t.cols = A; t.nc = f.ref_end; } // 444-451
    else { t.rows = A; t.nr = f.ref_end; t.cols = B; t.nc = f.query_end; }            // 452-459
    t.nr_full = t.nr;
    t.nc_full = t.nc;
    if (a.gap_min > 0) {
        const int32_t la = (int32_t)(a.ref_off[p + 1] - a.ref_off[p]);
        const int32_t lb = (int32_t)(a.qry_off[p + 1] - a.qry_off[p]);
        const bool fwd_rows_a = !(la < lb);
        const int32_t ec = fwd_rows_a ? f.query_end : f.ref_end;
        const int64_t span = (int64_t)ec + (int64_t)max(0, ubound - f.score) / a.gap_min;
        const int32_t lim = (int32_t)min<int64_t>(span, 0x7FFFFFFF);
        if (rev_rows_a == fwd_rows_a) t.nr = min(t.nr, lim);
        else t.nc = min(t.nc, lim);
    }
    return t;
}

// The reverse pass takes the other sequence as rows than the forward pass
// (kernel.cpp:248-259 decide by the full lengths, 444-459 by the end coordinates).
__device__ __forceinline__ bool reverse_swaps(const KernelArgs& a, int64_t p, const FwdOut& f)
{
    const int32_t la = (int32_t)(a.ref_off[p + 1] - a.ref_off[p]);
    const int32_t lb = (int32_t)(a.qry_off[p + 1] - a.qry_off[p]);
    return !(la < lb) != !(f.ref_end < f.query_end);
}

// The whole reversed prefix box A[0:ref_end) x B[0:query_end), no extent bound.
__device__ __forceinline__ Task reverse_task_full(const KernelArgs& a, int64_t p, const FwdOut& f)
{
    Task t;
    t.rev = 1;
    const uint8_t* A = a.ref + a.ref_off[p];
    const uint8_t* B = a.qry + a.qry_off[p];
    if (f.ref_end < f.query_end) { t.rows = B; t.nr = f.query_end; t.cols = A; t.nc = f.ref_end; } // 444-451
    else { t.rows = A; t.nr = f.ref_end; t.cols = B; t.nc = f.query_end; }                        // 452-459
    t.nr_full = t.nr;
    t.nc_full = t.nc;
    return t;
}

// Reverse target of a pair.  With a symmetric table, or an asymmetric one whose
// reverse pass keeps the forward orientation, the reversed prefix box holds the
// same alignments with the same scores, so its maximum is the forward score S
// and the extent bound applies.  With an asymmetric table and the orientation
// swapped, the reverse DP scores the transposed table and its maximum (which
// decides the begins, kernel.cpp:1021-1030, 3-45) is found first, by a
// forward-style pass over the whole reversed box (run by the kernels through
// their one forward-pass call site, so that no second copy of the pass is
// inlined).  A target of 0 gives begins (0, 0), the reference's initial
// per-thread coordinates.
struct RevPlan {
    Task t;
    int32_t target;
    bool gen;       // this group's pair needs the maximum of its reversed box first
    bool saturated; // that maximum may have saturated int16: recomputed with int32
};
__device__ __forceinline__ RevPlan reverse_plan(const KernelArgs& a, const Smem& sm, int64_t p, const FwdOut& f)
{
    RevPlan r;
    r.t = reverse_task(a, p, f, column_bound(a, sm, p, f));
    r.target = f.score;
    r.gen = a.asym && p >= 0 && f.score > 0 && reverse_swaps(a, p, f);
    r.saturated = false;
    if (r.gen) r.t = reverse_task_full(a, p, f);
    return r;
}
// The forward-style pass over the reversed box (r.t, or an empty task for
// groups that do not need it) returned `key`: its score is the target.
__device__ __forceinline__ void reverse_plan_finish(const KernelArgs& a, int64_t p, uint64_t key, RevPlan& r)
{
    if (!r.gen) return;
    r.target = (int32_t)(key >> 40);
    if (r.target <= 0) r.t = empty_task(a, 1);
    if (r.target >= a.sat_limit) {
        r.saturated = true;
        r.t = empty_task(a, 1);
        if ((threadIdx.x & 15) == 0) {
            const int32_t slot = atomicAdd(a.status + kStatusSaturated, 1);
            a.redo[slot] = (int32_t)p;
        }
    }
}

__device__ __forceinline__ void reverse_result(const KernelArgs& a, int64_t p, const FwdOut& f, const Task& t,
                                               int32_t target, uint64_t key, int32_t* ref_begin, int32_t* query_begin)
{
    *ref_begin = 0;
    *query_begin = 0;
    if (p < 0 || f.score <= 0 || target <= 0) return;
    if (key == 0) { atomicOr(a.status + kStatusRevMiss, 1); return; }
    const int32_t col = (int32_t)(0xFFFFFu - (uint32_t)((key >> 20) & 0xFFFFFu));
    const int32_t row = (int32_t)(0xFFFFFu - (uint32_t)(key & 0xFFFFFu));
    const int32_t orow = t.nr_full - 1 - row; // kernel.cpp:579
    const int32_t ocol = t.nc_full - 1 - col; // kernel.cpp:580
    if (f.ref_end < f.query_end) { *query_begin = orow; *ref_begin = ocol; } // 591-595
    else { *ref_begin = orow; *query_begin = ocol; }                       // 596-600
}

// Byte->code table, substitution scores and column maxima into LDS.
__device__ __forceinline__ void stage_tables(const KernelArgs& a, const Smem& sm)
{
    for (int i = threadIdx.x; i < 256; i += kWave) sm.lut()[i] = a.lut[i];
    for (int i = threadIdx.x; i < a.alpha * a.alpha; i += kWave) sm.sub()[i] = a.sub[i];
    for (int i = threadIdx.x; i < a.alpha; i += kWave) sm.colmax()[i] = a.colmax[i];
    __syncthreads();
}

__device__ __forceinline__ int64_t slot_pair(const KernelArgs& a, int64_t slot)
{
    if (slot >= a.order_n) return -1;
    return a.order ? (int64_t)a.order[slot] : slot;
}

// ---------------------------------------------------------------- kernels

template <class A, int K, bool P8>
__device__ __forceinline__ Smem make_smem(uint8_t* raw, const KernelArgs& a)
{
    LdsLayout L = lds_layout(prof_row_bytes(K, P8), a.alpha, (int)sizeof(typename A::B));
    // An opaque group offset: the compiler would otherwise split the rounding
    // constants of the layout into the profile address, leaving an odd byte offset
    // that ds_read2 cannot encode and one extra add per profile read.
    asm volatile("" : "+s"(L.group_off));
    return Smem{raw, L};
}

template <class A>
__device__ __forceinline__ typename A::B* group_bnd(const KernelArgs& a)
{
    const int g = threadIdx.x >> 4;
    return (typename A::B*)a.bnd + ((int64_t)blockIdx.x * kPairsPerWave + g) * a.bnd_rows;
}

// Plain form, forward pass: quads of consecutive pairs in input order, statically
// strided over the grid (one launch per batch, like driver.cpp:176-185).
template <int K, bool P8>
__global__ __launch_bounds__(64) void k_plain_forward(KernelArgs a)
{
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const Smem sm = make_smem<Packed16, K, P8>(smem_raw, a);
    stage_tables(a, sm);
    const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
    Packed16::B* bnd_g = group_bnd<Packed16>(a);
    const int64_t nquads = (a.order_n + kPairsPerWave - 1) / kPairsPerWave;
    for (int64_t q = blockIdx.x; q < nquads; q += gridDim.x) {
        const int64_t p = slot_pair(a, q * kPairsPerWave + g);
        const Task t = forward_task(a, p);
        const PassResult r = run_pass<Packed16, K, false, P8>(a, sm, t, 0, bnd_g);
        const FwdOut f = forward_result(a, p, r.key);
        if (p >= 0 && l == 0) {
            a.out[kOutScore * a.n + p] = f.score;
            a.out[kOutRefEnd * a.n + p] = f.ref_end;
            a.out[kOutQueryEnd * a.n + p] = f.query_end;
        }
    }
}

// Plain form, reverse pass: reads the forward results of the same buffer
// (driver.cpp:199-208, with no host round trip for the block size).
template <int K, bool P8>
__global__ __launch_bounds__(64) void k_plain_reverse(KernelArgs a)
{
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const Smem sm = make_smem<Packed16, K, P8>(smem_raw, a);
    stage_tables(a, sm);
    const int g = threadIdx.x >> 4, l = threadIdx.x & 15;
    Packed16::B* bnd_g = group_bnd<Packed16>(a);
    const int64_t nquads = (a.order_n + kPairsPerWave - 1) / kPairsPerWave;
    for (int64_t q = blockIdx.x; q < nquads; q += gridDim.x) {
        const int64_t p = slot_pair(a, q * kPairsPerWave + g);
        FwdOut f{0, 0, 0};
        if (p >= 0) {
            f.score = a.out[kOutScore * a.n + p];
            f.ref_end = a.out[kOutRefEnd * a.n + p];
            f.query_end = a.out[kOutQueryEnd * a.n + p];
        }
        RevPlan rp = reverse_plan(a, sm, p, f);
        if (__ballot(rp.gen) != 0) { // wave-uniform; groups without it idle on an empty task
            const PassResult rm = run_pass<Packed16, K, false, P8>(a, sm, rp.gen ? rp.t : empty_task(a, 1), 0, bnd_g);
            reverse_plan_finish(a, p, rm.key, rp);
        }
        const PassResult r = run_pass<Packed16, K, true, P8>(a, sm, rp.t, rp.target, bnd_g);
        int32_t rb, qb;
        reverse_result(a, p, f, rp.t, rp.target, r.key, &rb, &qb);
        if (p >= 0 && l == 0 && f.score >= 0 && !rp.saturated) {
            a.out[kOutRefBegin * a.n + p] = rb;
            a.out[kOutQueryBegin * a.n + p] = qb;
        }
    }
}

// Batched form: persistent workgroups pull quads of length-sorted pairs from a
// device work queue and run forward then reverse back to back.  With A = Wide32
// it recomputes the pairs the int16 kernels queued as possibly saturated.
template <class A, int K, bool P8>
__global__ __launch_bounds__(64) void k_batched(KernelArgs a)
{
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const Smem sm = make_smem<A, K, P8>(smem_raw, a);
    stage_tables(a, sm);
    const int l = threadIdx.x & 15;
    const int g = threadIdx.x >> 4;
    typename A::B* bnd_g = group_bnd<A>(a);
    const int64_t nquads = (a.order_n + kPairsPerWave - 1) / kPairsPerWave;
    for (;;) {
        uint32_t q = 0;
        if (threadIdx.x == 0) q = atomicAdd(a.queue, 1u);
        q = (uint32_t)__builtin_amdgcn_readfirstlane((int)q);
        if ((int64_t)q >= nquads) break;
        const int64_t p = slot_pair(a, (int64_t)q * kPairsPerWave + g);
        // the forward pass, then (asymmetric tables, wave-uniform) the
        // forward-style pass over the reversed boxes that need their own maximum:
        // one call site, so the pass is inlined once
        Task tf = forward_task(a, p);
        FwdOut f{0, 0, 0};
        RevPlan rp;
#pragma nounroll
        for (int fp = 0;; ++fp) {
            const PassResult rf = run_pass<A, K, false, P8>(a, sm, tf, 0, bnd_g);
            if (fp == 0) {
                f = forward_result(a, p, rf.key);
                rp = reverse_plan(a, sm, p, f);
                if (__ballot(rp.gen) == 0) break;
                tf = rp.gen ? rp.t : empty_task(a, 1);
            } else {
                reverse_plan_finish(a, p, rf.key, rp);
                break;
            }
        }
        const PassResult rr = run_pass<A, K, true, P8>(a, sm, rp.t, rp.target, bnd_g);
        int32_t rb, qb;
        reverse_result(a, p, f, rp.t, rp.target, rr.key, &rb, &qb);
        if (p >= 0 && l == 0 && f.score >= 0 && !rp.saturated) {
            a.out[kOutScore * a.n + p] = f.score;
            a.out[kOutRefEnd * a.n + p] = f.ref_end;
            a.out[kOutQueryEnd * a.n + p] = f.query_end;
            a.out[kOutRefBegin * a.n + p] = rb;
            a.out[kOutQueryBegin * a.n + p] = qb;
        }
    }
}

// Split form of k_batched for buckets of multi-strip pairs: a quad's strips run
// on different workgroups, pipelined, instead of one after another on one.
//
// Workgroup w serves XCD x = w % 8 (where the hardware places it on an 8-XCC
// partition), with its own work counter a.queue[x] and boundary slots, so
// producer and consumer of a boundary share an L2.  Each workgroup checks that
// assumption at entry against the hardware's XCC id (placement_consistent); a
// violation sets kStatusMisplaced and the host reruns the job unsplit.  XCD x owns quads x, x + 8, ...; its items are first
// (quad, strip) of the forward pass for all its quads, then the same for the
// reverse pass (a.split.strips items per quad and pass; strips past a quad's
// count are empty items).  Item i writes its outgoing boundary into slot
// i % slots (epoch i / slots) and reads the previous strip's from slot (i - 1),
// following that slot's progress word.  A slot is released once per epoch, in
// epoch order, by the strip that read it (or by its producer when no strip
// follows); a producer waits for the release of the epoch before.  Every wait is
// for an item taken from the counter earlier, which holds its workgroup until it
// finishes, so the waits end; each is bounded by kSpinTicks besides.
//
// Winner keys reduce with 64-bit atomicMax per pair.  A quad's reverse items
// start once its forward items are all done; the item completing the reverse
// count writes the quad's results.
//
// The waves take the highest issue priority: a split bucket is a chain of
// waits that runs beside the bulk buckets' waves on the same SIMDs, and every
// step it loses there delays the strips after it (the 1M-pair DNA mix: 31.18
// -> 30.98 ms per batch; protein sets, whose buckets are all split, unchanged).
// Its waits sleep between polls, so it takes no issue slots while waiting.
template <class A, int K, bool P8>
__global__ __launch_bounds__(64) void k_split(KernelArgs a)
{
    extern __shared__ __align__(16) uint8_t smem_raw[];
    const Smem sm = make_smem<A, K, P8>(smem_raw, a);
    stage_tables(a, sm);
    const int l = threadIdx.x & 15;
    const int g = threadIdx.x >> 4;
    const int x = (int)(blockIdx.x % kXcds);
    const SplitArgs& sp = a.split;
    // every workgroup of counter x must share one XCC (its L2): record this
    // one's, and flag the job for an unsplit rerun when they differ
    if (threadIdx.x == 0) {
        const uint32_t xcc = (uint32_t)__builtin_amdgcn_s_getreg(kXccIdReg) & 15u;
        const uint32_t prev = atomicCAS(sp.xcc + x, 0u, xcc + 1u);
        if (!placement_consistent(prev, xcc)) {
            // the results of this job are void: let every wait return at once
            atomicOr(a.status + kStatusMisplaced, 1);
            atomicOr(a.status + kStatusTimeout, 1);
        }
    }
    __builtin_amdgcn_s_setprio(3);
    const int64_t nquads = (a.order_n + kPairsPerWave - 1) / kPairsPerWave;
    const int64_t nq_x = nquads > x ? (nquads - x + kXcds - 1) / kXcds : 0;
    const int64_t per_pass = nq_x * sp.strips;
    typename A::B* bnd = (typename A::B*)a.bnd;
    const int64_t slot0 = (int64_t)x * sp.slots;
    auto slot_bnd = [&](int64_t i) {
        return bnd + ((slot0 + i % sp.slots) * kPairsPerWave + g) * a.bnd_rows;
    };
    // release epoch i / slots of slot i % slots, after the epoch before it
    auto release = [&](int64_t i) {
        uint32_t* r = sp.rel + slot0 + i % sp.slots;
        const uint32_t e = (uint32_t)(i / sp.slots);
        wait_u32(r, e, a.status);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        if (threadIdx.x == 0) __hip_atomic_store(r, e + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    for (;;) {
        uint32_t it = 0;
        if (threadIdx.x == 0) it = atomicAdd(a.queue + x, 1u);
        const int64_t i = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)it);
        if (i >= 2 * per_pass) break;
        const int rev = i >= per_pass;
        const int64_t k = rev ? i - per_pass : i;
        const int64_t q = (int64_t)x + (int64_t)kXcds * (k / sp.strips);
        const int s = (int)(k % sp.strips);
        const int64_t ws = q * kPairsPerWave + g; // work slot of this group's pair
        const int64_t p = slot_pair(a, ws);
        uint64_t* fkey = sp.keys;
        uint64_t* rkey = sp.keys + a.order_n;

        Task t;
        FwdOut f{0, 0, 0};
        if (!rev) {
            t = forward_task(a, p);
        } else {
            // the quad's forward strips have all finished
            wait_u32(sp.done + q, (uint32_t)sp.strips, a.status);
            const uint64_t fk = p >= 0 ? __hip_atomic_load(fkey + ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            f = forward_result(a, p, fk, s == 0);
            t = reverse_task(a, p, f, column_bound(a, sm, p, f));
            // A pair whose target an earlier strip reached is dropped from the
            // rest of its strips: they hold larger columns only.  The drop must be
            // monotone along the strips (a strip reads the pair's boundary rows,
            // and waits for the progress word, of the strip before), and two
            // strips read the winner key at different times, so strip s > 0 waits
            // for strip s - 1's decision in gate[ws], drops the pair if that
            // strip did, and publishes its own decision for strip s + 1.
            uint32_t* gw = p >= 0 ? sp.gate + ws : nullptr;
            const uint32_t prev = wait_group_u32(s > 0 ? gw : nullptr, (uint32_t)s, a.status);
            bool keep = t.nr > 0 && t.nc > 0 && prev != kGateDropped;
            if (keep && s > 0 && __hip_atomic_load(rkey + ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
                keep = false;
            if (!keep) t.nr = 0;
            if (gw && l == 0)
                __hip_atomic_store(gw, keep ? (uint32_t)s + 1 : kGateDropped, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
        }
        // strips of the quad (wave maximum)
        int st = (t.nr > 0 && t.nc > 0) ? (t.nc + kVirtPerPair * K - 1) / (kVirtPerPair * K) : 0;
#pragma unroll
        for (int o = 16; o < 64; o <<= 1) st = max(st, __shfl_xor(st, o));
        st = __builtin_amdgcn_readfirstlane(st);
        if (s < st) {
            wait_u32(sp.rel + slot0 + i % sp.slots, (uint32_t)(i / sp.slots), a.status); // slot free
            StripIO io;
            io.strip = s;
            io.bnd_in = s > 0 ? (const void*)slot_bnd(i - 1) : nullptr;
            io.bnd_out = (void*)slot_bnd(i);
            io.prog_in = s > 0 ? sp.prog + slot0 + (i - 1) % sp.slots : nullptr;
            io.prog_out = sp.prog + slot0 + i % sp.slots;
            io.epoch_in = s > 0 ? (uint64_t)((i - 1) / sp.slots) << 32 : 0;
            io.epoch_out = (uint64_t)(i / sp.slots) << 32;
            PassResult r;
            if (!rev) r = run_pass<A, K, false, P8, true>(a, sm, t, 0, nullptr, io);
            else r = run_pass<A, K, true, P8, true>(a, sm, t, f.score, nullptr, io);
            if (p >= 0 && l == 0 && r.key != 0) atomicMax((unsigned long long*)((rev ? rkey : fkey) + ws), r.key);
        }
        if (s > 0) release(i - 1);             // done with the previous strip's rows
        if (s == sp.strips - 1) release(i);    // no strip follows
        // count the item; the last reverse item of the quad writes its results
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
        uint32_t before = 0;
        if (threadIdx.x == 0) before = atomicAdd(sp.done + (rev ? nquads : 0) + q, 1u);
        before = (uint32_t)__builtin_amdgcn_readfirstlane((int)before);
        if (rev && before == (uint32_t)sp.strips - 1) {
            const uint64_t rk = p >= 0 ? __hip_atomic_load(rkey + ws, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0;
            int32_t rb, qb;
            reverse_result(a, p, f, t, f.score, rk, &rb, &qb); // (the host never splits asymmetric tables)
            if (p >= 0 && l == 0 && f.score >= 0) {
                a.out[kOutScore * a.n + p] = f.score;
                a.out[kOutRefEnd * a.n + p] = f.ref_end;
                a.out[kOutQueryEnd * a.n + p] = f.query_end;
                a.out[kOutRefBegin * a.n + p] = rb;
                a.out[kOutQueryBegin * a.n + p] = qb;
            }
        }
    }
}

__global__ void k_zero_u32(uint32_t* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- host side

template <int K, bool P8>
static const void* kernel_ptr(int kind)
{
    switch (kind) {
    case kPlainForward: return (const void*)&k_plain_forward<K, P8>;
    case kPlainReverse: return (const void*)&k_plain_reverse<K, P8>;
    case kBatched: return (const void*)&k_batched<Packed16, K, P8>;
    default: return nullptr;
    }
}

template <int K, bool P8>
static const void* wide_ptr()
{
    return (const void*)&k_batched<Wide32, K, P8>;
}

template <int K, bool P8>
static const void* split_ptr(bool f16)
{
    if (!f16) return (const void*)&k_split<Packed16, K, P8>;
    return (const void*)&k_split<PackedF16, K, P8>;
}

// (with P8, the byte profile holds the high bytes of f16 scores)
template <int K, bool P8>
static const void* f16_ptr()
{
    return (const void*)&k_batched<PackedF16, K, P8>;
}

// Compiled geometries (a build may narrow the lists with -D for inspection).
#ifndef GBSW_K_LIST
#define GBSW_K_LIST(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16) X(20) X(24) X(28) X(32)
#endif

#ifndef GBSW_WIDE_K_LIST
#define GBSW_WIDE_K_LIST(X) X(4) X(8) X(16)
#endif

// f16 kernels: the geometries whose single strip (32 x K columns) keeps DNA-like
// scores below 2048
#ifndef GBSW_F16_K_LIST
#define GBSW_F16_K_LIST(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16) X(20)
#endif

// split kernels: the geometries multi-strip buckets use
#ifndef GBSW_SPLIT_K_LIST
#define GBSW_SPLIT_K_LIST(X) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(10) X(12) X(14) X(16)
#endif

template <bool P8>
static const void* kernel_for_profile(int kind, int K)
{
    if (kind == kBatchedWide) {
        switch (K) {
#define GBSW_CASE(k) case k: return wide_ptr<k, P8>();
            GBSW_WIDE_K_LIST(GBSW_CASE)
#undef GBSW_CASE
        default: return nullptr;
        }
    }
    if (kind == kBatchedF16) {
        switch (K) {
#define GBSW_CASE(k) case k: return f16_ptr<k, P8>();
            GBSW_F16_K_LIST(GBSW_CASE)
#undef GBSW_CASE
        default: return nullptr;
        }
    }
    if (kind == kSplit || kind == kSplitF16) {
        switch (K) {
#define GBSW_CASE(k) case k: return split_ptr<k, P8>(kind == kSplitF16);
            GBSW_SPLIT_K_LIST(GBSW_CASE)
#undef GBSW_CASE
        default: return nullptr;
        }
    }
    switch (K) {
#define GBSW_CASE(k) case k: return kernel_ptr<k, P8>(kind);
        GBSW_K_LIST(GBSW_CASE)
#undef GBSW_CASE
    default: return nullptr;
    }
}

// kind may carry kByteProfile
static const void* kernel_for(int kind, int K)
{
    return (kind & kByteProfile) ? kernel_for_profile<true>(kind & ~kByteProfile, K)
                                 : kernel_for_profile<false>(kind, K);
}

bool kernel_supported(int K) { return kernel_for(kPlainForward, K) != nullptr; }
bool kernel_has(int kind, int K) { return kernel_for(kind, K) != nullptr; }

const int* wide_k_list(int* count)
{
    static const int ks[] = {
#define GBSW_ITEM(k) k,
        GBSW_WIDE_K_LIST(GBSW_ITEM)
#undef GBSW_ITEM
    };
    *count = (int)(sizeof(ks) / sizeof(ks[0]));
    return ks;
}

const int* kernel_k_list(int* count)
{
    static const int ks[] = {
#define GBSW_ITEM(k) k,
        GBSW_K_LIST(GBSW_ITEM)
#undef GBSW_ITEM
    };
    *count = (int)(sizeof(ks) / sizeof(ks[0]));
    return ks;
}

// Workgroups of one kernel that fit on a CU at this LDS size (0 on error).
int kernel_occupancy(int kind, int K, int lds_bytes)
{
    const void* f = kernel_for(kind, K);
    if (!f) return 0;
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, kWave, (size_t)lds_bytes) != hipSuccess) return 0;
    return nb;
}

hipError_t kernel_launch(int kind, int K, int grid, int lds_bytes, hipStream_t stream, const KernelArgs& args)
{
    const void* f = kernel_for(kind, K);
    if (!f) return hipErrorInvalidValue;
    if (lds_bytes > 65536) {
        const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        if (e != hipSuccess) return e;
    }
    KernelArgs a = args;
    void* params[] = {&a};
    return hipLaunchKernel(f, dim3(grid), dim3(kWave), params, (size_t)lds_bytes, stream);
}

hipError_t zero_u32(uint32_t* p, int n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

} // namespace gbsw
