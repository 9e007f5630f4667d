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
//  k_tb_dp:      one wave per pair (four per workgroup).  Lane l owns column 64 s + l + 1 of stripe s
//                and walks the rectangle's rows along anti-diagonals (row
//                t - l + 1 at step t); the left neighbour's H and E arrive through
//                a one-lane DPP shift, the stripe's left boundary column through
//                memory.  Each cell stores one byte of tightness bits (which of
//                its predecessors reach its H, E and F values), laid out step-major
//                per stripe, so a step's 64 bytes are one coalesced store.
//  k_tb_walk:    one lane per pair: walks the bits back from the rectangle's end
//                corner, choosing the smallest op that an optimal alignment can
//                continue with, and writes run-length ops, last run first.
//  k_tb_compact: one lane per pair: copies its runs, reversed, to their final
//                place.
#include <hip/hip_runtime.h>

#include "gbsw_kernels.hpp"

namespace gbsw {

namespace {

constexpr int32_t kTbNeg = -(1 << 29);

enum : uint32_t {
    kHD = 1u,      // H == diagonal + s
    kHF = 2u,      // H == F (H == E when neither of these holds)
    kFExt = 8u,    // F == F(up) + extend
    kFOpen = 16u,  // F == H(up) + open
    kEExt = 32u,   // E == E(left) + extend
    kEOpen = 64u,  // E == H(left) + open
};

enum : uint32_t { kOpM = 0, kOpI = 1, kOpD = 2 }; // BAM CIGAR op codes

// value of the lane one below (wave_shr:1); lane 0 receives 0 (bound_ctrl)
__device__ __forceinline__ int32_t from_left(int32_t v)
{
    return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);
}

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

__host__ __device__ __forceinline__ int64_t tb_dir_bytes(int32_t m, int32_t n)
{
    const int64_t stripes = (n + 63) / 64;
    return (stripes * (int64_t)(m + 63) * 64 + 7) & ~(int64_t)7;
}

constexpr int kTbWaves = 4;    // pairs (waves) per workgroup of k_tb_dp
constexpr int kTbBlock = 256;  // steps between refills of a wave's row ring
constexpr int kTbRing = 512;   // row codes per wave: a block's rows and the 63 before

__global__ __launch_bounds__(64 * kTbWaves) void k_tb_dp(TbArgs a, int64_t count)
{
    __shared__ uint8_t lut[256];
    __shared__ int16_t sub[kMaxAlpha * kMaxAlpha];
    __shared__ uint8_t ring_all[kTbWaves][kTbRing];
    __shared__ uint64_t bndbuf_all[kTbWaves][kTbBlock];
    for (int x = threadIdx.x; x < 256; x += 64 * kTbWaves) lut[x] = a.lut[x];
    for (int x = threadIdx.x; x < a.alpha * a.alpha; x += 64 * kTbWaves) sub[x] = a.sub[x];
    __syncthreads();
    // one wave per pair from here on: no workgroup barriers below

    // the wave index in a scalar register: the pair's sizes and addresses stay scalar
    const int l = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t k = (int64_t)blockIdx.x * kTbWaves + wv;
    if (k >= count) return;
    const int64_t p = a.order[a.p0 + k];
    if (a.out[p] <= 0) {
        if (l == 0) a.path_score[p] = 0;
        return;
    }
    uint8_t* ring = ring_all[wv];
    uint64_t* bndbuf = bndbuf_all[wv];
    const PairRect r = pair_rect(a, p);
    const int32_t m = r.m, n = r.n, alpha = a.alpha;
    const int32_t go = a.gap_open, ge = a.gap_extend;
    uint8_t* dir = a.dir + a.dir_off[k];
    uint64_t* bnd = (uint64_t*)(dir + tb_dir_bytes(m, n)); // boundary column, rows 1..m
    const int stripes = (n + 63) / 64;
    for (int s = 0; s < stripes; ++s) {
        const int32_t c = 64 * s + l + 1;
        const bool vc = c <= n;
        const int cb = vc ? lut[r.B[c - 1]] : 0;
        int32_t Hup = go + (c - 1) * ge, Fup = kTbNeg; // row 0
        int32_t Hout = 0, Eout = kTbNeg, Hdiag = 0;
        // byte offset of this lane's cell at step t: dcol + 64 t (a pair's region
        // is below 4 GiB)
        const uint32_t dcol = (uint32_t)((int64_t)s * (m + 63) * 64) + (uint32_t)l;
        for (int32_t t0 = 0; t0 < m + 63; t0 += kTbBlock) {
            // rows [t0, t0 + 256) enter the ring as codes (rows t0 - 63 .. t0 - 1
            // stay from the previous block), with the left boundary of those rows;
            // lane l reads row t - l at step t, lane 0 the boundary of row t + 1
            for (int q = 0; q < kTbBlock / 64; ++q) {
                const int32_t x = t0 + q * 64 + l;
                if (x < m) ring[x & (kTbRing - 1)] = lut[r.A[x]];
                if (s > 0 && x + 1 <= m)
                    bndbuf[q * 64 + l] = __hip_atomic_load(bnd + x + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_s_waitcnt(0); // this wave's ring writes land before its reads
            const int32_t t1 = min(t0 + kTbBlock, m + 63);
            for (int32_t t = t0; t < t1; ++t) {
                const int32_t i = t - l + 1;
                // left boundary of lane 0's row t + 1 (a uniform branch and a
                // broadcast read; rows past the rectangle read unused entries)
                int32_t bH, bE;
                if (s == 0) {
                    bH = go + t * ge; // H(t + 1, 0): a gap from the corner
                    bE = kTbNeg;
                } else {
                    const uint64_t v = bndbuf[t - t0];
                    bH = (int32_t)(uint32_t)v;
                    bE = (int32_t)(uint32_t)(v >> 32);
                }
                // the shifts run in every lane (a DPP read from a lane outside the
                // exec mask would return the fill value)
                const int32_t sH = from_left(Hout), sE = from_left(Eout);
                const int32_t Hl = l == 0 ? bH : sH;
                const int32_t El = l == 0 ? bE : sE;
                if (i >= 1 && i <= m && vc) {
                    const int32_t diag = i == 1 ? (c == 1 ? 0 : go + (c - 2) * ge) : Hdiag;
                    const int code = ring[(t - l) & (kTbRing - 1)];
                    const int32_t sv = r.swap ? sub[cb * alpha + code] : sub[code * alpha + cb];
                    const int32_t E = max(El + ge, Hl + go);
                    const int32_t F = max(Fup + ge, Hup + go);
                    const int32_t D = diag + sv;
                    const int32_t H = max(D, max(E, F));
                    const uint32_t bits = (H == D ? kHD : 0u) | (H == F ? kHF : 0u) |
                                          (F == Fup + ge ? kFExt : 0u) | (F == Hup + go ? kFOpen : 0u) |
                                          (E == El + ge ? kEExt : 0u) | (E == Hl + go ? kEOpen : 0u);
                    dir[dcol + (uint32_t)t * 64u] = (uint8_t)bits;
                    Hup = H;
                    Fup = F;
                    Hout = H;
                    Eout = E;
                    if (l == 63 && s + 1 < stripes) bnd[i] = (uint64_t)(uint32_t)H | ((uint64_t)(uint32_t)E << 32);
                    if (i == m && c == n) a.path_score[p] = H;
                }
                Hdiag = Hl;
            }
        }
        // the boundary column is in memory before the next stripe's agent-scope
        // loads read it (one wave: its own stores, complete)
        __builtin_amdgcn_s_waitcnt(0);
    }
}

__device__ __forceinline__ uint32_t tb_bits(const uint8_t* dir, int32_t m, int32_t i, int32_t j)
{
    const int32_t s = (j - 1) >> 6, l = (j - 1) & 63;
    return dir[((int64_t)s * (m + 63) + (i + l - 1)) * 64 + l];
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
    const int64_t bnd = n > 64 ? 8 * ((int64_t)m + 1) : 0;
    return tb_dir_bytes(m, n) + bnd;
}

hipError_t tb_launch_dp(const TbArgs& a, int64_t count, hipStream_t stream)
{
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tb_dp, dim3((unsigned)((count + kTbWaves - 1) / kTbWaves)), dim3(64 * kTbWaves), 0, stream, a,
                       count);
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
