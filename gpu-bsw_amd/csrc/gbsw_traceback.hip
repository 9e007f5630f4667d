// gbsw_traceback.hip -- alignment paths (CIGAR) of aligned pairs on gfx950.
//
// The reference declares a traceback (traceBack, kernel.hpp:36-39) but never
// defines it.  Here the path of a pair with score > 0 is the best GLOBAL
// alignment of ref[ref_begin, ref_end) against query[query_begin, query_end)
// (the rectangle the forward and reverse passes report), scored like the forward
// pass: a run of L gap positions scores open + (L-1) extend, and the substitution
// is s(row residue, column residue) with rows = the longer full sequence (ref on
// ties, kernel.cpp:248-259).  Among optimal alignments the path is the one whose
// op sequence, read from its end, is lexicographically smallest with M < D < I.
//
// Three kernels per call:
//  k_tb_dp16 / k_tb_dp (int16 / int32 cells; the host takes the int16 form when
//                every path value of the batch fits it):
//  k_tb_dp:      four pairs per wave, one per 16-lane row ("group").  Lane l of a
//                group owns columns 64 s + 4 l + 1 .. + 4 of stripe s and walks the
//                rectangle's rows along anti-diagonals (row t - l + 1 at step t);
//                the left neighbour's H and E arrive through a one-lane DPP shift
//                inside the row, the stripe's left boundary column through memory,
//                the substitution scores of the lane's four columns from a per-stripe
//                profile in LDS (one 8-byte read per step).  Each cell gets six
//                tightness bits (which of its predecessors reach its H, E and F
//                values); a lane's four cells of a step share one dword, a group's
//                sixteen dwords are one coalesced 64-byte store.
//  k_tb_dp16:    the same with two pairs per lane, one in each 16-bit half of every
//                state register (eight pairs per wave): each v_pk_* op advances
//                two cells, and a bit costs three packed ops for two cells.
//  k_tb_walk:    one lane per pair: walks the bits back from the rectangle's end
//                corner, choosing the smallest op that an optimal alignment can
//                continue with, and writes run-length ops, last run first.
//  k_tb_compact: one lane per pair: copies its runs, reversed, to their final
//                place.
#include <hip/hip_runtime.h>

#include "gbsw_device.hpp"
#include "gbsw_kernels.hpp"

namespace gbsw {

namespace {

constexpr int32_t kTbNeg = -(1 << 29);

// Tightness of a cell: six bits, a lane's four cells of a step in one dword
// (columns 0 and 1 of the lane in bits 6..11 and 0..5, columns 2 and 3 in bits
// 22..27 and 16..21, so that either half of the dword is one 16-bit accumulator).
// Stored, most significant first: the E field (2 bits: 0 when only E(left) +
// extend reaches E, 2 when only H(left) + open does, 1 when both), the F field
// (likewise, from above), then H == F and H == D; tb_bits turns them into these
// flags.
enum : uint32_t {
    kHD = 1u,      // H == diagonal + s
    kHF = 2u,      // H == F (H == E when neither of these holds)
    kFExt = 4u,    // F == F(up) + extend
    kFOpen = 8u,   // F == H(up) + open
    kEExt = 16u,   // E == E(left) + extend
    kEOpen = 32u,  // E == H(left) + open
};

enum : uint32_t { kOpM = 0, kOpI = 1, kOpD = 2 }; // BAM CIGAR op codes

struct PairRect {
    int32_t m, n;    // rectangle rows (ref side) and columns (query side)
    bool swap;       // rows of the forward pass are the query (|ref| < |query|)
    const uint8_t* A;
    const uint8_t* B;
};

__device__ __forceinline__ PairRect pair_rect(const TbArgs& a, int64_t p)
{
    PairRect r;
    const int32_t rb = a.out[a.n + p], re = a.out[2 * a.n + p];
    const int32_t qb = a.out[3 * a.n + p], qe = a.out[4 * a.n + p];
    r.m = re - rb;
    r.n = qe - qb;
    r.swap = (a.ref_off[p + 1] - a.ref_off[p]) < (a.qry_off[p + 1] - a.qry_off[p]);
    r.A = a.ref + a.ref_off[p] + rb;
    r.B = a.qry + a.qry_off[p] + qb;
    return r;
}

constexpr int kTbLanes = 16;              // lanes per pair
constexpr int kTbK = 4;                   // columns per lane
constexpr int kTbCols = kTbStripeColumns; // columns per stripe (kTbLanes x kTbK)
static_assert(kTbCols == kTbLanes * kTbK, "stripe width");
constexpr int kTbLag = kTbLanes - 1;      // steps by which the last lane trails the first
constexpr int kTbGroups = 64 / kTbLanes;  // pairs per wave (workgroup)
constexpr int kTbBlock = 64;              // steps between refills of a group's row ring
constexpr int kTbRing = 128;              // row codes per group: a block's rows and the 15 before

// Bytes of a pair's tightness bits: per stripe, (rows + 15) steps x 16 lanes x 4 columns.
__host__ __device__ __forceinline__ int64_t tb_dir_bytes(int32_t m, int32_t n)
{
    const int64_t stripes = (n + kTbCols - 1) / kTbCols;
    return (stripes * (int64_t)(m + kTbLag) * kTbCols + 7) & ~(int64_t)7;
}

// LDS of k_tb_dp: byte -> code table, alpha^2 scores, then per group the profile
// ((alpha + 1) rows x 16 lanes x 4 int16), the row ring (codes as profile row
// offsets) and the block's boundary entries.
struct TbLds {
    int32_t sub_off, group_off, group_bytes, ring_off, bnd_off, total;
};
__host__ __device__ __forceinline__ TbLds tb_lds(int32_t alpha)
{
    TbLds L;
    L.sub_off = 256;
    L.group_off = (L.sub_off + 2 * alpha * alpha + 15) & ~15;
    const int32_t prof = (alpha + 1) * kTbLanes * kTbK * 2;
    L.ring_off = prof;
    L.bnd_off = (L.ring_off + 2 * kTbRing + 15) & ~15;
    L.group_bytes = L.bnd_off + 8 * kTbBlock;
    L.total = L.group_off + kTbGroups * L.group_bytes;
    return L;
}

// value of the lane one below inside the 16-lane row (row_shr:1); lane 0 of the
// row receives 0 (bound_ctrl; every group's first lane takes the boundary instead)
__device__ __forceinline__ int32_t from_left16(int32_t v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x111, 0xF, 0xF, true);
}

// 2 acc + (x == y): the compare's VCC is the carry-in of a doubling add (two
// instructions per bit; the compiler's own form is a compare, a select and a merge)
__device__ __forceinline__ uint32_t shift_in_eq(uint32_t acc, int32_t x, int32_t y)
{
    asm("v_cmp_eq_u32 vcc, %1, %2\n\tv_addc_co_u32 %0, vcc, %0, %0, vcc" : "+v"(acc) : "v"(x), "v"(y) : "vcc");
    return acc;
}

__global__ __launch_bounds__(64) void k_tb_dp(TbArgs a, int64_t count)
{
    extern __shared__ __align__(16) uint8_t tb_smem[];
    const TbLds L = tb_lds(a.alpha);
    uint8_t* lut = tb_smem;
    int16_t* sub = (int16_t*)(tb_smem + L.sub_off);
    for (int x = threadIdx.x; x < 256; x += 64) lut[x] = a.lut[x];
    for (int x = threadIdx.x; x < a.alpha * a.alpha; x += 64) sub[x] = a.sub[x];
    __syncthreads();

    const int lane = threadIdx.x, g = lane / kTbLanes, l = lane % kTbLanes;
    uint8_t* gb = tb_smem + L.group_off + g * L.group_bytes;
    int16_t* prof = (int16_t*)gb;
    uint16_t* ring = (uint16_t*)(gb + L.ring_off);
    uint64_t* bndbuf = (uint64_t*)(gb + L.bnd_off);
    const int32_t alpha = a.alpha, go = a.gap_open, ge = a.gap_extend;

    // this group's pair (none: an idle group that computes nothing it stores)
    const int64_t k = (int64_t)blockIdx.x * kTbGroups + g;
    int64_t p = -1;
    if (k < count) {
        p = a.order[a.p0 + k];
        if (a.out[p] <= 0) {
            if (l == 0) a.path_score[p] = 0;
            p = -1;
        }
    }
    PairRect r{0, 0, false, nullptr, nullptr};
    if (p >= 0) r = pair_rect(a, p);
    const int32_t m = r.m, n = r.n;
    const int stripes_g = (p >= 0) ? (n + kTbCols - 1) / kTbCols : 0;
    uint8_t* dir = (p >= 0) ? a.dir + a.dir_off[k] : nullptr;
    uint64_t* bnd = (p >= 0) ? (uint64_t*)(dir + tb_dir_bytes(m, n)) : nullptr; // boundary column, rows 1..m
    // wave-uniform extents: the largest stripe and step counts of the four groups
    int stripes = stripes_g, rows = stripes_g ? m : 0;
#pragma unroll
    for (int o = kTbLanes; o < 64; o <<= 1) {
        stripes = max(stripes, __shfl_xor(stripes, o));
        rows = max(rows, __shfl_xor(rows, o));
    }
    stripes = __builtin_amdgcn_readfirstlane(stripes);
    rows = __builtin_amdgcn_readfirstlane(rows);
    const int32_t steps = rows + kTbLag;

    for (int s = 0; s < stripes; ++s) {
        const bool live = s < stripes_g;
        const int32_t c0 = kTbCols * s + kTbK * l + 1; // the lane's first column (1-based)
        // profile of the stripe: row a, lane l, column k = s(code a, column c0 + k)
        // in the pair's orientation (row code a indexes the rectangle's rows)
        {
            int cb[kTbK];
#pragma unroll
            for (int q = 0; q < kTbK; ++q) {
                const int32_t c = c0 + q;
                cb[q] = (live && c <= n) ? lut[r.B[c - 1]] : -1;
                cb[q] = cb[q] >= alpha ? -1 : cb[q];
            }
            for (int row = 0; row <= alpha; ++row) {
                uint32_t w[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int32_t v2[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int c = cb[2 * h + e];
                        v2[e] = (row < alpha && c >= 0) ? (r.swap ? sub[c * alpha + row] : sub[row * alpha + c]) : 0;
                    }
                    w[h] = (uint32_t)(uint16_t)v2[0] | ((uint32_t)(uint16_t)v2[1] << 16);
                }
                *(uint2*)(prof + (row * kTbLanes + l) * kTbK) = make_uint2(w[0], w[1]);
            }
        }
        // row 0 of the rectangle: H(0, c) = open + (c - 1) extend, F = -inf
        int32_t Hup[kTbK], Fup[kTbK];
#pragma unroll
        for (int q = 0; q < kTbK; ++q) {
            Hup[q] = go + (c0 + q - 1) * ge;
            Fup[q] = kTbNeg;
        }
        int32_t Hout = 0, Eout = kTbNeg;
        int32_t Hdiag = 0; // H(row - 1, c0 - 1), set from the left input of the previous step
        const uint32_t dbase = (uint32_t)((int64_t)s * (m + kTbLag) * kTbCols) + (uint32_t)(4 * l);
        for (int32_t t0 = 0; t0 < steps; t0 += kTbBlock) {
            // rows t0 + 1 .. t0 + 64 enter the ring as profile offsets (the 15 rows
            // before stay from the previous block), and lane 0's boundary entries
            // of those rows (stripe s - 1 is complete: this wave wrote them)
#pragma unroll
            for (int q = 0; q < kTbBlock / kTbLanes; ++q) {
                const int32_t x = t0 + q * kTbLanes + l; // row x + 1
                int code = alpha;
                if (live && x < m) {
                    code = lut[r.A[x]];
                    code = code > alpha ? alpha : code;
                }
                ring[x & (kTbRing - 1)] = (uint16_t)(code * kTbLanes * kTbK * 2);
                if (live && s > 0 && x < m)
                    bndbuf[q * kTbLanes + l] = __hip_atomic_load(bnd + x + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_s_waitcnt(0); // this wave's ring writes land before its reads
            const int32_t t1 = min(t0 + kTbBlock, steps);
            for (int32_t t = t0; t < t1; ++t) {
                const int32_t i = t - l + 1; // row of this lane
                // left inputs: lane l - 1's last column of the previous step; lane 0
                // takes the left boundary of row t + 1
                int32_t bH, bE;
                if (s == 0) {
                    bH = go + t * ge; // H(t + 1, 0): a gap from the corner
                    bE = kTbNeg;
                } else {
                    const uint64_t v = bndbuf[t - t0];
                    bH = (int32_t)(uint32_t)v;
                    bE = (int32_t)(uint32_t)(v >> 32);
                }
                // (the shifts run in every lane: a DPP read from a lane outside the
                // exec mask would return the fill value)
                const int32_t sH = from_left16(Hout), sE = from_left16(Eout);
                const int32_t Hl = l == 0 ? bH : sH;
                const int32_t El = l == 0 ? bE : sE;
                const int32_t Hd = Hdiag; // H(i - 1, c0 - 1)
                Hdiag = Hl;
                // a lane's state changes only on its rows of the rectangle; its outputs
                // on other rows reach only its right neighbour's rows outside it
                if (live && i >= 1 && i <= m) {
                    // the lane's four scores of row i
                    const uint2 pw = *(const uint2*)((const uint8_t*)prof + ring[(i - 1) & (kTbRing - 1)] + l * kTbK * 2);
                    const int32_t sc[kTbK] = {(int32_t)(int16_t)pw.x, (int32_t)pw.x >> 16, (int32_t)(int16_t)pw.y,
                                              (int32_t)pw.y >> 16};
                    int32_t diag = i == 1 ? (c0 == 1 ? 0 : go + (c0 - 2) * ge) : Hd;
                    int32_t Hleft = Hl, Eleft = El;
                    uint32_t word = 0, word01 = 0;
#pragma unroll
                    for (int q = 0; q < kTbK; ++q) {
                        const int32_t E = max(Eleft + ge, Hleft + go);
                        const int32_t F = max(Fup[q] + ge, Hup[q] + go);
                        const int32_t D = diag + sc[q];
                        const int32_t H = max(D, max(E, F));
                        // the E and F fields, then the two H bits
                        word = (word << 2) | (uint32_t)(1 + (E == Hleft + go) - (E == Eleft + ge));
                        word = (word << 2) | (uint32_t)(1 + (F == Hup[q] + go) - (F == Fup[q] + ge));
                        word = shift_in_eq(word, H, F); // kHF
                        word = shift_in_eq(word, H, D); // kHD
                        diag = Hup[q]; // H(i - 1, column q): the next column's diagonal
                        Hup[q] = H;
                        Fup[q] = F;
                        Hleft = H;
                        Eleft = E;
                        if (q == 1) {
                            word01 = word;
                            word = 0;
                        }
                    }
                    Hout = Hleft;
                    Eout = Eleft;
                    *(uint32_t*)(dir + dbase + (uint32_t)t * kTbCols) = word01 | (word << 16);
                    if (l == kTbLanes - 1 && s + 1 < stripes_g)
                        bnd[i] = (uint64_t)(uint32_t)Hout | ((uint64_t)(uint32_t)Eout << 32);
                    if (i == m && n >= c0 && n < c0 + kTbK) {
                        int32_t hn = 0;
#pragma unroll
                        for (int q = 0; q < kTbK; ++q)
                            if (c0 + q == n) hn = Hup[q];
                        a.path_score[p] = hn;
                    }
                }
            }
        }
        // the boundary column is in memory before the next stripe's agent-scope
        // loads read it (one wave: its own stores, complete)
        __builtin_amdgcn_s_waitcnt(0);
    }
}

// ---- k_tb_dp16: two pairs per lane, one per 16-bit half

constexpr int kTb16Pairs = 2 * kTbGroups; // pairs per wave

// LDS of k_tb_dp16: byte -> code table, alpha^2 scores, then per group the two
// pairs' profiles, the row ring (both pairs' profile offsets in one dword), the
// block's incoming boundary entries and its outgoing ones (Ho and E of both pairs,
// packed).
struct TbLds16 {
    int32_t sub_off, group_off, prof_bytes, ring_off, bnd_off, bout_off, group_bytes, total;
};
__host__ __device__ __forceinline__ TbLds16 tb_lds16(int32_t alpha)
{
    TbLds16 L;
    L.sub_off = 256;
    L.group_off = (L.sub_off + 2 * alpha * alpha + 15) & ~15;
    L.prof_bytes = (alpha + 1) * kTbLanes * kTbK * 2;
    L.ring_off = 2 * L.prof_bytes;
    L.bnd_off = (L.ring_off + 4 * kTbRing + 15) & ~15;
    L.bout_off = L.bnd_off + 8 * kTbBlock;
    L.group_bytes = L.bout_off + 8 * kTbBlock;
    L.total = L.group_off + kTbGroups * L.group_bytes;
    return L;
}

__device__ __forceinline__ uint32_t splat16(int32_t v)
{
    const uint32_t h = (uint32_t)(uint16_t)(int16_t)v;
    return h | (h << 16);
}
__device__ __forceinline__ int32_t half16(uint32_t x, int h)
{
    return (int32_t)(int16_t)(uint16_t)(h ? x >> 16 : x & 0xFFFFu);
}
// A cell's six tightness bits shifted into each half of acc (modulo 2^16 per half,
// v_pk_mad_u16, so a field written as its value - 1 borrows only inside its half;
// tb_fields_bias adds the ones back), for E = max(eo, ee), F = max(fo, fe) and
// H = max(D, E, F):
//   field - 1 of E (of F) = the sign of eo - ee: v_pk_sub_i16 clamp, then
//     v_pk_min_i16 / v_pk_max_i16 to [-1, 1];
//   H == F (H == D): v_pk_sub_u16 gives H - F >= 0, v_pk_sub_u16 clamp turns it
//     into 1 - (H - F) floored at 0, 1 exactly when equal.
// (Written as asm: the compiler's own form of the clamp was two compares and two
// selects per half, and of the multiply-add a shift and an add.)
__device__ __forceinline__ uint32_t tb_cell_bits16(uint32_t acc, uint32_t eo, uint32_t ee, uint32_t fo, uint32_t fe,
                                                   uint32_t H, uint32_t F, uint32_t D)
{
    uint32_t t0, t1;
    asm("v_pk_sub_i16 %1, %3, %4 clamp\n\t"
        "v_pk_sub_i16 %2, %5, %6 clamp\n\t"
        "v_pk_min_i16 %1, %1, 1 op_sel_hi:[1,0]\n\t"
        "v_pk_min_i16 %2, %2, 1 op_sel_hi:[1,0]\n\t"
        "v_pk_max_i16 %1, %1, -1\n\t"
        "v_pk_max_i16 %2, %2, -1\n\t"
        "v_pk_mad_u16 %0, %0, 4, %1 op_sel_hi:[1,0,1]\n\t"
        "v_pk_mad_u16 %0, %0, 4, %2 op_sel_hi:[1,0,1]\n\t"
        "v_pk_sub_u16 %1, %7, %8\n\t"
        "v_pk_sub_u16 %2, %7, %9\n\t"
        "v_pk_sub_u16 %1, 1, %1 op_sel_hi:[0,1] clamp\n\t"
        "v_pk_sub_u16 %2, 1, %2 op_sel_hi:[0,1] clamp\n\t"
        "v_pk_mad_u16 %0, %0, 2, %1 op_sel_hi:[1,0,1]\n\t"
        "v_pk_mad_u16 %0, %0, 2, %2 op_sel_hi:[1,0,1]"
        : "+v"(acc), "=&v"(t0), "=&v"(t1)
        : "v"(eo), "v"(ee), "v"(fo), "v"(fe), "v"(H), "v"(F), "v"(D));
    return acc;
}
// the ones of the four fields of an accumulator half (two columns: the E and F
// fields end at bits 4 and 2 of a column's six, the first column six higher)
constexpr uint32_t kTbFieldOnes = (1u << 10) | (1u << 8) | (1u << 4) | (1u << 2);
__device__ __forceinline__ uint32_t tb_fields_bias(uint32_t acc)
{
    return raw(as_u(acc) + (u16x2){(uint16_t)kTbFieldOnes, (uint16_t)kTbFieldOnes});
}

// The state holds H + open (Ho) in place of H: both open candidates of a cell are
// then the neighbours' values as they are, and the profile holds s - open so that
// the diagonal Ho(i-1, j-1) + s - open is D itself.  (H + open is a value of the
// same range.)
// Cells are int16 with saturating adds: the host takes this kernel when
// (rows + columns + 4) x the largest |score| of every pair stays below 32,000, so
// every path value and every sum formed from one stays inside the int16 range
// and only the -inf entries (-32768) saturate; they equal no cell value, as the
// int32 kernel's -2^29 does not.  The bits are those of k_tb_dp, in the same
// layout; the boundary column (read only by this kernel's next stripe) holds Ho
// and E.
__global__ __launch_bounds__(64) void k_tb_dp16(TbArgs a, int64_t count)
{
    extern __shared__ __align__(16) uint8_t tb_smem[];
    const TbLds16 L = tb_lds16(a.alpha);
    uint8_t* lut = tb_smem;
    int16_t* sub = (int16_t*)(tb_smem + L.sub_off);
    for (int x = threadIdx.x; x < 256; x += 64) lut[x] = a.lut[x];
    for (int x = threadIdx.x; x < a.alpha * a.alpha; x += 64) sub[x] = a.sub[x];
    __syncthreads();

    const int lane = threadIdx.x, g = lane / kTbLanes, l = lane % kTbLanes;
    uint8_t* gb = tb_smem + L.group_off + g * L.group_bytes;
    uint32_t* ring = (uint32_t*)(gb + L.ring_off);
    uint2* bndbuf = (uint2*)(gb + L.bnd_off);
    uint2* bout = (uint2*)(gb + L.bout_off);
    const int32_t alpha = a.alpha, go = a.gap_open, ge = a.gap_extend;

    // the group's two pairs, h = 0 in the low halves (none: an idle half whose
    // cells are computed and never stored)
    int64_t p[2];
    PairRect r[2];
    int32_t m[2], n[2];
    int stripes_h[2];
    uint8_t* dir[2];
    uint64_t* bnd[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int64_t k = (int64_t)blockIdx.x * kTb16Pairs + 2 * g + h;
        p[h] = -1;
        if (k < count) {
            p[h] = a.order[a.p0 + k];
            if (a.out[p[h]] <= 0) {
                if (l == 0) a.path_score[p[h]] = 0;
                p[h] = -1;
            }
        }
        r[h] = PairRect{0, 0, false, nullptr, nullptr};
        if (p[h] >= 0) r[h] = pair_rect(a, p[h]);
        m[h] = r[h].m;
        n[h] = r[h].n;
        stripes_h[h] = (p[h] >= 0) ? (n[h] + kTbCols - 1) / kTbCols : 0;
        dir[h] = (p[h] >= 0) ? a.dir + a.dir_off[k] : nullptr;
        bnd[h] = (p[h] >= 0) ? (uint64_t*)(dir[h] + tb_dir_bytes(m[h], n[h])) : nullptr;
    }
    // wave-uniform extents: the largest stripe and row counts of the eight pairs
    int stripes = max(stripes_h[0], stripes_h[1]);
    int rows = max(stripes_h[0] ? m[0] : 0, stripes_h[1] ? m[1] : 0);
#pragma unroll
    for (int o = kTbLanes; o < 64; o <<= 1) {
        stripes = max(stripes, __shfl_xor(stripes, o));
        rows = max(rows, __shfl_xor(rows, o));
    }
    stripes = __builtin_amdgcn_readfirstlane(stripes);
    rows = __builtin_amdgcn_readfirstlane(rows);
    const int32_t steps = rows + kTbLag;
    const uint32_t GO2 = splat16(go), GE2 = splat16(ge), NEG2 = 0x80008000u;

    for (int s = 0; s < stripes; ++s) {
        const bool live[2] = {s < stripes_h[0], s < stripes_h[1]};
        const int32_t mlive = max(live[0] ? m[0] : 0, live[1] ? m[1] : 0);
        const int32_t c0 = kTbCols * s + kTbK * l + 1; // the lane's first column (1-based)
        // profiles of the stripe (one per pair, as in k_tb_dp)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            int16_t* prof = (int16_t*)(gb + h * L.prof_bytes);
            int cb[kTbK];
#pragma unroll
            for (int q = 0; q < kTbK; ++q) {
                const int32_t c = c0 + q;
                cb[q] = (live[h] && c <= n[h]) ? lut[r[h].B[c - 1]] : -1;
                cb[q] = cb[q] >= alpha ? -1 : cb[q];
            }
            for (int row = 0; row <= alpha; ++row) {
                uint32_t w[2];
#pragma unroll
                for (int e2 = 0; e2 < 2; ++e2) {
                    int32_t v2[2];
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        const int c = cb[2 * e2 + e];
                        v2[e] = ((row < alpha && c >= 0) ? (r[h].swap ? sub[c * alpha + row] : sub[row * alpha + c]) : 0) - go;
                    }
                    w[e2] = (uint32_t)(uint16_t)v2[0] | ((uint32_t)(uint16_t)v2[1] << 16);
                }
                *(uint2*)(prof + (row * kTbLanes + l) * kTbK) = make_uint2(w[0], w[1]);
            }
        }
        // row 0 of the rectangle: H(0, c) = open + (c - 1) extend, F = -inf
        uint32_t Hup[kTbK], Fup[kTbK]; // Ho and F of the row above
#pragma unroll
        for (int q = 0; q < kTbK; ++q) {
            Hup[q] = splat16(2 * go + (c0 + q - 1) * ge);
            Fup[q] = NEG2;
        }
        uint32_t Hout = splat16(go), Eout = NEG2; // Ho and E of the lane's last column
        uint32_t Hdiag = 0; // Ho(row - 1, c0 - 1), set from the left input of the previous step
        const uint32_t diag1 = splat16((c0 == 1 ? 0 : go + (c0 - 2) * ge) + go); // Ho(0, c0 - 1)
        // the step at which this lane meets a pair's end corner (row m, column n), -1: none
        int32_t tpath[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            tpath[h] = (live[h] && n[h] >= c0 && n[h] < c0 + kTbK) ? m[h] + l - 1 : -1;
        uint32_t dbase[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
            dbase[h] = (uint32_t)((int64_t)s * (m[h] + kTbLag) * kTbCols) + (uint32_t)(4 * l);
        for (int32_t t0 = 0; t0 < steps; t0 += kTbBlock) {
            // rows t0 + 1 .. t0 + 64 of both pairs enter the ring (profile offsets)
            // and, after the first stripe, lane 0's boundary entries of those rows
#pragma unroll
            for (int q = 0; q < kTbBlock / kTbLanes; ++q) {
                const int32_t x = t0 + q * kTbLanes + l; // row x + 1
                uint32_t off[2];
                uint64_t v[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    int code = alpha;
                    v[h] = 0;
                    if (live[h] && x < m[h]) {
                        code = lut[r[h].A[x]];
                        code = code > alpha ? alpha : code;
                        if (s > 0)
                            v[h] = __hip_atomic_load(bnd[h] + x + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    }
                    off[h] = (uint32_t)(code * kTbLanes * kTbK * 2);
                }
                ring[x & (kTbRing - 1)] = off[0] | (off[1] << 16);
                if (s > 0)
                    bndbuf[q * kTbLanes + l] =
                        make_uint2(__builtin_amdgcn_perm((uint32_t)v[1], (uint32_t)v[0], 0x05040100u),
                                   __builtin_amdgcn_perm((uint32_t)(v[1] >> 32), (uint32_t)(v[0] >> 32), 0x05040100u));
            }
            __builtin_amdgcn_s_waitcnt(0); // this wave's ring writes land before its reads
            const int32_t t1 = min(t0 + kTbBlock, steps);
            for (int32_t t = t0; t < t1; ++t) {
                const int32_t i = t - l + 1; // row of this lane
                uint32_t bH, bE;
                if (s == 0) {
                    bH = splat16(2 * go + t * ge); // Ho(t + 1, 0): a gap from the corner
                    bE = NEG2;
                } else {
                    const uint2 v = bndbuf[t - t0];
                    bH = v.x;
                    bE = v.y;
                }
                const uint32_t sH = (uint32_t)from_left16((int32_t)Hout), sE = (uint32_t)from_left16((int32_t)Eout);
                const uint32_t Hl = l == 0 ? bH : sH;
                const uint32_t El = l == 0 ? bE : sE;
                const uint32_t Hd = Hdiag; // Ho(i - 1, c0 - 1)
                Hdiag = Hl;
                if (i >= 1 && i <= mlive) {
                    // the lane's four scores of row i, pair 0 low and pair 1 high
                    const uint32_t rw = ring[(i - 1) & (kTbRing - 1)];
                    const uint2 pa = *(const uint2*)(gb + (rw & 0xFFFFu) + l * kTbK * 2);
                    const uint2 pb = *(const uint2*)(gb + L.prof_bytes + (rw >> 16) + l * kTbK * 2);
                    const uint32_t sc[kTbK] = {__builtin_amdgcn_perm(pb.x, pa.x, 0x05040100u),
                                               __builtin_amdgcn_perm(pb.x, pa.x, 0x07060302u),
                                               __builtin_amdgcn_perm(pb.y, pa.y, 0x05040100u),
                                               __builtin_amdgcn_perm(pb.y, pa.y, 0x07060302u)};
                    uint32_t diag = i == 1 ? diag1 : Hd;
                    uint32_t Hleft = Hl, Eleft = El;
                    uint32_t word = 0, word01 = 0;
#pragma unroll
                    for (int q = 0; q < kTbK; ++q) {
                        const uint32_t eo = Hleft, ee = pk_add_sat(Eleft, GE2);
                        const uint32_t fo = Hup[q], fe = pk_add_sat(Fup[q], GE2);
                        const uint32_t E = pk_max(eo, ee);
                        const uint32_t F = pk_max(fo, fe);
                        const uint32_t D = pk_add_sat(diag, sc[q]);
                        const uint32_t H = pk_max(D, pk_max(E, F));
                        // the E and F fields, then H == F and H == D
                        word = tb_cell_bits16(word, eo, ee, fo, fe, H, F, D);
                        const uint32_t Ho = pk_add_sat(H, GO2);
                        diag = Hup[q];
                        Hup[q] = Ho;
                        Fup[q] = F;
                        Hleft = Ho;
                        Eleft = E;
                        if (q == 1) {
                            word01 = word;
                            word = 0;
                        }
                    }
                    Hout = Hleft;
                    Eout = Eleft;
                    word01 = tb_fields_bias(word01);
                    word = tb_fields_bias(word);
                    // each pair's dword: its halves of the two accumulators
                    const uint32_t dw[2] = {__builtin_amdgcn_perm(word, word01, 0x05040100u),
                                            __builtin_amdgcn_perm(word, word01, 0x07060302u)};
                    // the last lane's column goes to the next stripe: staged here,
                    // stored a block at a time
                    if (l == kTbLanes - 1) bout[t - t0] = make_uint2(Hout, Eout);
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        if (live[h] && i <= m[h]) {
                            *(uint32_t*)(dir[h] + dbase[h] + (uint32_t)t * kTbCols) = dw[h];
                            if (t == tpath[h]) {
                                int32_t hn = 0;
#pragma unroll
                                for (int q = 0; q < kTbK; ++q)
                                    if (c0 + q == n[h]) hn = half16(Hup[q], h) - go;
                                a.path_score[p[h]] = hn;
                            }
                        }
                    }
                }
            }
            // the block's outgoing boundary entries: step t0 + e holds row t0 + e - 14
            // of the last lane (written by this group's lane 15; the same wave, so
            // its LDS writes have landed once the wait returns)
            __builtin_amdgcn_s_waitcnt(0);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (s + 1 < stripes_h[h]) {
                    for (int e = l; e < t1 - t0; e += kTbLanes) {
                        const int32_t row = t0 + e - (kTbLanes - 2);
                        if (row >= 1 && row <= m[h]) {
                            const uint2 v = bout[e];
                            bnd[h][row] = (uint64_t)(uint32_t)half16(v.x, h) | ((uint64_t)(uint32_t)half16(v.y, h) << 32);
                        }
                    }
                }
            }
        }
        // the boundary columns are in memory before the next stripe's agent-scope
        // loads read them (one wave: its own stores, complete)
        __builtin_amdgcn_s_waitcnt(0);
    }
}

// bits of cell (i, j): stripe s, lane l, column q of the lane; step i - 1 + l
__device__ __forceinline__ uint32_t tb_bits(const uint8_t* dir, int32_t m, int32_t i, int32_t j)
{
    const int32_t s = (j - 1) / kTbCols, c = (j - 1) % kTbCols, l = c / kTbK, q = c % kTbK;
    const uint32_t w = *(const uint32_t*)(dir + ((int64_t)s * (m + kTbLag) + (i - 1 + l)) * kTbCols + kTbK * l);
    const uint32_t b = (w >> (16 * (q >> 1) + 6 * (1 - (q & 1)))) & 63u;
    const uint32_t ef = b >> 4, ff = (b >> 2) & 3u;
    return (b & (kHD | kHF)) | (ef >= 1 ? kEOpen : 0u) | (ef <= 1 ? kEExt : 0u) | (ff >= 1 ? kFOpen : 0u) |
           (ff <= 1 ? kFExt : 0u);
}

__global__ __launch_bounds__(64) void k_tb_walk(TbArgs a, int64_t count)
{
    const int64_t k = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (k >= count) return;
    const int64_t p = a.order[a.p0 + k];
    if (a.out[p] <= 0) {
        a.nruns[p] = 0;
        return;
    }
    const PairRect r = pair_rect(a, p);
    const int32_t m = r.m;
    const uint8_t* dir = a.dir + a.dir_off[k];
    uint32_t* runs = a.runs + a.cap_off[p];
    int32_t nr = 0;
    uint32_t run_op = 0, run_len = 0;
    auto emit = [&](uint32_t op, uint32_t len) {
        if (len == 0) return;
        if (run_len && op == run_op) {
            run_len += len;
        } else {
            if (run_len) runs[nr++] = run_len << 4 | run_op;
            run_op = op;
            run_len = len;
        }
    };
    int32_t i = m, j = r.n, st = 0; // 0: H, 1: in a D run (F), 2: in an I run (E)
    while (i > 0 && j > 0) {
        const uint32_t b = tb_bits(dir, m, i, j);
        if (st == 0) {
            if (b & kHD) {
                emit(kOpM, 1);
                --i;
                --j;
            } else {
                st = (b & kHF) ? 1 : 2;
            }
        } else if (st == 1) {
            emit(kOpD, 1);
            const bool m_next = (b & kFOpen) && i >= 2 && (tb_bits(dir, m, i - 1, j) & kHD);
            st = m_next ? 0 : (b & kFExt) ? 1 : 0;
            --i;
        } else {
            emit(kOpI, 1);
            const bool md_next = (b & kEOpen) && j >= 2 && (tb_bits(dir, m, i, j - 1) & (kHD | kHF));
            st = md_next ? 0 : (b & kEExt) ? 2 : 0;
            --j;
        }
    }
    emit(kOpI, (uint32_t)j); // a border reached: the rest is one gap
    emit(kOpD, (uint32_t)i);
    if (run_len) runs[nr++] = run_len << 4 | run_op;
    a.nruns[p] = nr;
}

__global__ __launch_bounds__(256) void k_tb_compact(const uint32_t* runs, const int64_t* cap_off,
                                                    const int32_t* nruns, const int64_t* dst_off, uint32_t* dst,
                                                    int64_t n)
{
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int32_t c = nruns[p];
    const uint32_t* src = runs + cap_off[p];
    uint32_t* d = dst + dst_off[p];
    for (int32_t x = 0; x < c; ++x) d[x] = src[c - 1 - x];
}

} // namespace

int64_t tb_region_bytes(int32_t m, int32_t n)
{
    if (m <= 0 || n <= 0) return 0;
    const int64_t bnd = n > kTbCols ? 8 * ((int64_t)m + 1) : 0;
    return tb_dir_bytes(m, n) + bnd;
}

hipError_t tb_launch_dp(const TbArgs& a, int64_t count, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    if (a.cells16) {
        hipLaunchKernelGGL(k_tb_dp16, dim3((unsigned)((count + kTb16Pairs - 1) / kTb16Pairs)), dim3(64),
                           (size_t)tb_lds16(a.alpha).total, stream, a, count);
    } else {
        hipLaunchKernelGGL(k_tb_dp, dim3((unsigned)((count + kTbGroups - 1) / kTbGroups)), dim3(64),
                           (size_t)tb_lds(a.alpha).total, stream, a, count);
    }
    return hipGetLastError();
}

hipError_t tb_launch_walk(const TbArgs& a, int64_t count, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tb_walk, dim3((unsigned)((count + 63) / 64)), dim3(64), 0, stream, a, count);
    return hipGetLastError();
}

hipError_t tb_compact(const uint32_t* runs, const int64_t* cap_off, const int32_t* nruns, const int64_t* dst_off,
                      uint32_t* dst, int64_t n, hipStream_t stream)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tb_compact, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, runs, cap_off, nruns,
                       dst_off, dst, n);
    return hipGetLastError();
}

} // namespace gbsw
