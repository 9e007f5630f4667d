// gbsw_kernel_args.hpp -- argument block and LDS layout shared by host and kernels.
#pragma once
#include <stdint.h>

// The pure helpers below also build in host-only C++ (CPU tests).
#if !defined(__HIPCC__) && !defined(__host__)
#define __host__
#define __device__
#endif

namespace gbsw {

constexpr int kLanesPerPair = 16;              // one DPP row
constexpr int kVirtPerPair = 2 * kLanesPerPair; // two int16 halves per lane
constexpr int kPairsPerWave = 4;
constexpr int kWave = 64;
constexpr int kRing = 128;   // row-stream ring entries per pair
constexpr int kChunk = 32;   // rows refilled at a time
// The row-code ring keeps a copy of its first kRingTail entries after its end,
// so the entries a virtual lane reads during one chunk (32 steps, one step more
// when the step count is odd, and the look-ahead of one step) are contiguous
// whatever the chunk's position in the ring, also when read 16 entries on from
// the position of the virtual lane 16 higher (both halves of a lane use one
// pointer).
constexpr int kRingTail = kChunk + kLanesPerPair + 2;
constexpr int kRingCodes = kRing + kRingTail;
// Boundary-column ring: the look-ahead read of the step after a chunk's last
// may pass the ring's end by one or two entries (never used).
constexpr int kRingBnd = kRing + 2;
constexpr int kMaxAlpha = 31; // codes 0..A-1, row A is the padding row
constexpr int16_t kNeg = -16384;

// Output layout: five int32 arrays of n entries, back to back.
enum { kOutScore = 0, kOutRefBegin = 1, kOutRefEnd = 2, kOutQueryBegin = 3, kOutQueryEnd = 4 };

// status[] words written by kernels
enum {
    kStatusBadResidue = 0,
    kStatusSaturated = 1,
    kStatusRevMiss = 2,
    kStatusTimeout = 3,
    kStatusMisplaced = 4, // k_split: workgroups of one XCD counter ran on different XCCs
    kStatusWords = 5
};

// Split form (k_split): the strips of a multi-strip pair run on different
// workgroups of one XCD, each strip following its predecessor's boundary rows.
struct SplitArgs {
    uint64_t* keys;   // 2 x order_n: forward, then reverse winner keys by work slot
    uint32_t* done;   // 2 x quads: forward, then reverse strip items finished
    uint64_t* prog;   // per boundary slot: epoch << 32 | boundary rows ready
    uint32_t* rel;    // per boundary slot: epochs released
    uint32_t* xcc;    // kXcds words: 1 + the hardware XCC id seen by counter x's workgroups (0: none yet)
    uint32_t* gate;   // order_n words by work slot: last reverse strip's decision for the pair
                      // (s + 1: strip s keeps it, kGateDropped: dropped from the rest)
    int32_t strips;   // strip items per quad and pass
    int32_t slots;    // boundary slots per XCD
};
constexpr uint32_t kGateDropped = 0xFFFFFFFFu;
constexpr int kXcds = 8;                 // workgroup w is expected on XCD w % 8
constexpr uint64_t kSpinTicks = 200000000ull; // 2 s of the 100 MHz s_memrealtime clock

// Placement check of the split form.  Its producer and consumer strips exchange
// boundary rows through plain stores, which are visible to each other only
// inside one XCC's L2, so every workgroup serving counter x must run on one XCC.
// Each workgroup compare-and-swaps (1 + its XCC id) into slot x of a zeroed
// table; `prev` is what the slot held before.  The placement is consistent when
// the slot was empty or already held this XCC.  Whatever the hardware's
// workgroup -> XCC mapping (w % 8 on an 8-XCC partition, fewer XCCs per
// partition in other modes), only a mapping that puts two workgroups of one
// counter on different XCCs fails, and then the host reruns the job unsplit.
__host__ __device__ constexpr bool placement_consistent(uint32_t prev, uint32_t xcc)
{
    return prev == 0u || prev == xcc + 1u;
}
// The host's decision after a job: rerun it with the split form off.
__host__ __device__ constexpr bool split_rerun_needed(const int32_t* status)
{
    return status[kStatusMisplaced] != 0;
}
// s_getreg_b32 operand of HW_REG_XCC_ID (register 20, bits 3:0): the XCC the
// wave runs on (gfx942 / gfx950).
constexpr int kXccIdReg = (3 << 11) | (0 << 6) | 20;

struct KernelArgs {
    const uint8_t* ref;
    const uint64_t* ref_off;   // n+1 exclusive prefix offsets
    const uint8_t* qry;
    const uint64_t* qry_off;
    int64_t n;                 // pairs in the batch
    const int32_t* order;      // pair index of each work slot (nullptr = identity)
    int64_t order_n;           // slots in this launch
    uint32_t* queue;           // persistent work counter (batched)
    int32_t* out;              // 5*n int32
    void* bnd;                 // strip boundary rows, kPairsPerWave*bnd_rows entries per workgroup
    int64_t bnd_rows;
    int32_t* status;
    int32_t* redo;             // pairs whose int16 score may have saturated (count: status[kStatusSaturated])
    const uint8_t* lut;        // byte -> residue code, 0xFF for bytes outside the alphabet
    const int16_t* sub;        // A*A scores, index row_code*A + col_code
    const int16_t* colmax;     // per column code: max(0, max_a sub[a*A + code])
    int32_t alpha;             // A
    int32_t gap_min;           // min(-gap_open, -gap_extend): cost floor of one gap position
    int32_t gap_open;          // <= 0
    int32_t gap_extend;        // <= 0
    int32_t sat_limit;         // scores >= this may have saturated int16
    int32_t asym;              // 1: sub[] is not symmetric (the reverse pass then finds its own maximum
                               //    when its orientation differs from the forward pass's)
    SplitArgs split;           // k_split only
};

// Dynamic LDS layout of one 64-lane workgroup; W = profile columns per row
// (32 virtual lanes x kstride(K)).  The candidate exchange of a pass's final
// reduction runs after every group has left its sweep, so it reuses the start
// of the group areas.
struct LdsLayout {
    int32_t lut_off, sub_off, colmax_off, xch_off, info_off, group_off, group_bytes;
    int32_t prof_bytes, ring_code_off, ring_bnd_off, total;
};

__host__ __device__ inline int32_t align_up(int32_t x, int32_t a) { return (x + a - 1) / a * a; }

// Profile layout of one row for K columns per virtual lane: K/8 blocks of 8
// columns for all 32 virtual lanes (one b128 read each), then the K%8 remaining
// columns for all virtual lanes padded to a b32/b64/b96/b128 read (kremstride).
// Six columns rather than eight for K%8 = 5, 6 keep the LDS of a DNA workgroup
// at K = 5, 6 small enough for 12 workgroups per CU instead of 11.
__host__ __device__ constexpr int32_t kremstride(int32_t K)
{
    return K % 8 == 0 ? 0 : K % 8 <= 2 ? 2 : K % 8 <= 4 ? 4 : K % 8 <= 6 ? 6 : 8;
}
// Profile columns reserved per virtual lane.
__host__ __device__ constexpr int32_t kstride(int32_t K) { return K / 8 * 8 + kremstride(K); }

// Byte profile (alphabets of kByteProfileAlpha codes or more, i.e. protein,
// with every score in [-127, 127]): one int8 per column, -128 for padding, so
// the (alpha+1)-row profile of a strip takes half the LDS and leaves room for
// more waves per CU.  Layout of one row: K/16 blocks of 16 columns for all 32
// virtual lanes (one b128 read each), then the K%16 remaining columns padded to
// a b32/b64/b96/b128 read (kremstride8).  With the f16 kernels (DNA-like
// tables whose scores are f16 values with a zero low byte) a byte is the high
// byte of the score's f16 bits instead.
constexpr int kByteProfileAlpha = 8;
__host__ __device__ constexpr int32_t kremstride8(int32_t K)
{
    return K % 16 == 0 ? 0 : K % 16 <= 4 ? 4 : K % 16 <= 8 ? 8 : K % 16 <= 12 ? 12 : 16;
}
__host__ __device__ constexpr int32_t kstride8(int32_t K) { return K / 16 * 16 + kremstride8(K); }

// Bytes of one profile row (one residue code, all columns of a strip).
__host__ __device__ constexpr int32_t prof_row_bytes(int32_t K, bool bytes8)
{
    return bytes8 ? kVirtPerPair * kstride8(K) : kVirtPerPair * kstride(K) * 2;
}

// row_bytes: prof_row_bytes of the kernel; bnd_bytes: bytes of one
// strip-boundary entry (4 for packed int16, 8 for int32).
__host__ __device__ inline LdsLayout lds_layout(int32_t row_bytes, int32_t alpha, int32_t bnd_bytes)
{
    LdsLayout L;
    L.lut_off = 0;                                       // 256 B
    L.sub_off = 256;                                     // alpha^2 int16
    L.colmax_off = L.sub_off + align_up(2 * alpha * alpha, 16); // 32 int16
    L.info_off = L.colmax_off + 2 * 32;                  // 4 groups x 8 int32
    L.group_off = align_up(L.info_off + 4 * 8 * 4, 128);
    L.xch_off = L.group_off;                             // 64 lanes x 2 keys x 8 B
    L.prof_bytes = align_up((alpha + 1) * row_bytes, 128);
    L.ring_code_off = L.prof_bytes;
    L.ring_bnd_off = L.ring_code_off + kRingCodes * 2;
    L.group_bytes = align_up(L.ring_bnd_off + kRingBnd * bnd_bytes, 128);
    L.total = L.group_off + kPairsPerWave * L.group_bytes;
    return L;
}

} // namespace gbsw
