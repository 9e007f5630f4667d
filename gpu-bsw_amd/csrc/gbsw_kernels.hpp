// gbsw_kernels.hpp -- host-visible entry points of gbsw_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "gbsw_kernel_args.hpp"

namespace gbsw {

enum KernelKind {
    kPlainForward = 0,
    kPlainReverse = 1,
    kBatched = 2,
    kBatchedWide = 3,
    kBatchedF16 = 4,
    kSplit = 5,    // k_split, int16 cells
    kSplitF16 = 6, // k_split, f16 cells
};
// Flag on a kind: the instantiation with an int8 profile (no f16 form).
constexpr int kByteProfile = 0x100;
inline int with_profile(int kind, bool bytes8) { return bytes8 ? (kind | kByteProfile) : kind; }

// largest score + one substitution an f16 kernel may meet (exact integers in f16)
constexpr int kF16Exact = 2048;

bool kernel_supported(int K);
bool kernel_has(int kind, int K);
const int* kernel_k_list(int* count); // K values with compiled kernels, ascending
const int* wide_k_list(int* count);   // K values of the int32 (kBatchedWide) kernel
int kernel_occupancy(int kind, int K, int lds_bytes);
hipError_t kernel_launch(int kind, int K, int grid, int lds_bytes, hipStream_t stream, const KernelArgs& args);
hipError_t zero_u32(uint32_t* p, int n, hipStream_t stream);

// gbsw_binning.hip
// Columns of one strip (32 x K) of each bucket id.
struct StripWidths {
    uint16_t w[64];
};
hipError_t bin_keys(const uint64_t* ref_off, const uint64_t* qry_off, int64_t n, const uint8_t* bucket_of_len,
                    const StripWidths& widths, uint32_t* keys, int32_t* vals, hipStream_t stream);
// OR into mask[8] the set of byte values in p[0, n) (p 16-byte aligned).
hipError_t byte_presence(const uint8_t* p, uint64_t n, uint32_t* mask, hipStream_t stream);
hipError_t sort_pairs(void* temp, size_t& temp_bytes, const uint32_t* kin, uint32_t* kout, const int32_t* vin,
                      int32_t* vout, int n, int end_bit, hipStream_t stream);

// gbsw_gather.hip: one tile = 1..kGatherTileBytes bytes of one request part,
// copied from src to dst; part_flags[part] |= 1 when a byte has accept[byte] == 0.
constexpr uint32_t kGatherTileBytes = 32768;
struct GatherTile {
    const uint8_t* src; // device-readable pinned host memory
    uint8_t* dst;       // device memory
    uint32_t len;
    int32_t part;
};
hipError_t gather(const GatherTile* tiles, int ntiles, const uint8_t* accept, int32_t* part_flags,
                  hipStream_t stream);


// gbsw_traceback.hip: alignment paths of aligned pairs.
struct TbArgs {
    const uint8_t* ref;
    const uint8_t* qry;
    const uint64_t* ref_off;
    const uint64_t* qry_off;
    const int32_t* out;     // results, 5 x n (score, ref_begin, ref_end, query_begin, query_end)
    int64_t n;
    const uint8_t* lut;     // byte -> residue code
    const int16_t* sub;     // alpha x alpha, s(row code, column code)
    int32_t alpha;
    int32_t gap_open, gap_extend;
    int64_t p0;             // first position of the launch in `order`
    const int64_t* order;   // pairs by position (longest first)
    const int64_t* dir_off; // per position of the launch: byte offset of its region in dir
    uint8_t* dir;           // tightness bytes and boundary columns of the launch's pairs
    const int64_t* cap_off; // per pair: offset of its run capacity (rows + columns) in runs
    uint32_t* runs;         // run-length ops, last run first
    int32_t* nruns;         // per pair: runs written
    int32_t* path_score;    // per pair: score of the path
    int32_t cells16;        // 1: every path value fits int16 (k_tb_dp16), 0: int32 cells (k_tb_dp)
};
// Columns of one stripe of the traceback DP.
constexpr int kTbStripeColumns = 64;
// Scratch bytes of a pair with an m x n rectangle (0 when empty).
int64_t tb_region_bytes(int32_t m, int32_t n);
// Tightness bytes of the pairs at positions p0 .. p0 + count - 1 of `order`.
hipError_t tb_launch_dp(const TbArgs& a, int64_t count, hipStream_t stream);
// Their paths (run-length ops, last run first), from the bytes tb_launch_dp wrote.
hipError_t tb_launch_walk(const TbArgs& a, int64_t count, hipStream_t stream);
// dst[dst_off[p] ...] = runs of pair p in path order.
hipError_t tb_compact(const uint32_t* runs, const int64_t* cap_off, const int32_t* nruns, const int64_t* dst_off,
                      uint32_t* dst, int64_t n, hipStream_t stream);

} // namespace gbsw
