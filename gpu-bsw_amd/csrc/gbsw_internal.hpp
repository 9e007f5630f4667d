// gbsw_internal.hpp -- host-side building blocks shared by the resident-batch
// API (gbsw_api.cpp) and the serving loop (gbsw_server.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gbsw.h"
#include "gbsw_kernels.hpp"

namespace gbsw {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define GBSW_HIP_TRY(expr)                                                                         \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return ::gbsw::fail(GBSW_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Device memory owned by the buffer (grown by reserve), or a view of memory
// owned elsewhere (view(): reserve never allocates and fails beyond its size).
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    bool owned = true;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release();
    hipError_t reserve(size_t b);
    void view(void* p, size_t b)
    {
        release();
        owned = false;
        ptr = p;
        bytes = b;
    }
    template <class T> T* as() const { return (T*)ptr; }
};

struct PinnedBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    PinnedBuffer() = default;
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
    ~PinnedBuffer() { release(); }
    void release();
    hipError_t reserve(size_t b);
    template <class T> T* as() const { return (T*)ptr; }
};

// One length bucket of the batched form: a kernel geometry and a slice of the
// sorted work order.
struct Bucket {
    int K = 0;
    int64_t begin = 0, count = 0;
    int64_t cells = 0;
    int32_t max_cols = 0; // longest shorter sequence: bounds the bucket's scores
};

struct LaunchRec {
    int kind, K, grid;
    int64_t pairs, cells;
    double ms;
};

constexpr int kAuxStreams = 7;   // at most 1 + 7 bucket streams
constexpr int kDefaultStreams = 4;

// Streams, events and scratch of one in-flight alignment job.
struct Exec {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[kAuxStreams] = {};
    hipEvent_t ev_fork = nullptr, ev_join[kAuxStreams] = {};
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<hipEvent_t> kev; // per-launch event pairs
    DeviceBuffer bnd, queue, status, sub, lut, colmax, redo, split_buf;
    PinnedBuffer status_host;    // status words copied back at the end of a job
    PinnedBuffer tables_host;    // staging of sub / lut / colmax for the async copies
    std::vector<LaunchRec> launches;
    int32_t max_sub = 0;
    bool p8 = false;               // the current job uses the int8 profile
    int streams = kDefaultStreams; // streams for the length buckets (1 = serialised)
    // creates the auxiliary streams the first n streams need
    int ensure_streams(int n);
    bool f16 = true;               // exact f16 kernels where the score bound allows
    bool split = true;             // multi-strip buckets run their strips on several workgroups
    bool split_misplaced = false;  // a split launch ran workgroups of one XCD counter on different
                                   // XCCs: the split form stays off for this Exec
    bool late_results = false;     // last complete_job wrote results after the job's own copies
                                   // (int32 recompute or unsplit rerun)
    bool fixed = false;            // buffers preallocated (serving slot, or views of the
                                   // allocation-free scratch): never grown by a job; grids
                                   // shrink to fit its boundary rows
    bool graphs = true;            // batched form of a resident batch replays a captured hipGraph
    int graph_state = 0;           // last job: 0 eager, 1 captured, 2 replayed
    int kmodel = 1;                // length -> geometry model of new batches (kModel*)
    int init(int device, int num_cu);
    // streams, events and pinned staging only; every device buffer is a view set
    // by the caller (the allocation-free entry point), launch events pre-created
    int init_views(int device, int num_cu);
    void destroy();
};

// Device view of the pairs of one job (a resident batch, or a serving slot).
// The captured stream work of a batch's batched form, with what it depends on.
struct JobGraph {
    hipGraphExec_t exec = nullptr;
    std::string key;
    std::vector<LaunchRec> launches; // records of the capture
    int64_t captures = 0, replays = 0;
    void reset()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        exec = nullptr;
        key.clear();
    }
    JobGraph() = default;
    JobGraph(const JobGraph&) = delete;
    JobGraph& operator=(const JobGraph&) = delete;
    ~JobGraph() { reset(); }
};

struct Job {
    JobGraph* graph = nullptr;       // set for resident batches: graph capture allowed
    int32_t alphabet = GBSW_DNA;
    int64_t n = 0;
    const uint8_t* ref = nullptr;
    const uint8_t* qry = nullptr;
    const uint64_t* ref_off = nullptr;
    const uint64_t* qry_off = nullptr;
    int32_t* out = nullptr;          // 5*n int32
    int32_t alpha = 0;               // residue codes
    const uint8_t* code_of = nullptr; // protein: reference code (0..23) of each residue code;
                                      // nullptr: residue code k is reference code k (alpha 24)
    int kmodel = 1;                  // length -> geometry model (kModel*), fixed at creation
    const uint8_t* lut_host = nullptr; // 256 entries
    int32_t max_rows = 0, max_cols = 0;
    int64_t cells = 0;
    std::vector<Bucket> buckets;     // batched form (host counts)
    const uint8_t* bucket_of_len = nullptr; // device table for the keys kernel
    uint32_t* keys = nullptr;
    uint32_t* keys_sorted = nullptr;
    int32_t* vals = nullptr;
    int32_t* order = nullptr;
    void* sort_temp = nullptr;
    size_t sort_temp_bytes = 0;
};

// Largest compiled K that fits a 64 KiB workgroup for this alphabet size and
// profile width.
int max_k_for(int alpha, bool p8);
// Large alphabets use the int8 profile when the scoring fits it; geometries of a
// batch are chosen assuming it.
inline bool byte_profile_alphabet(int alpha) { return alpha >= kByteProfileAlpha; }
// Length -> geometry models: strips x (K + 2) minimised, or that cost divided by
// the VALU issue share the geometry's occupancy allows.
constexpr int kModelClassic = 0;
constexpr int kModelOccupancy = 1;
// models 1.. divide the cost by min(1, waves per SIMD / threshold), threshold
// 2.5, 2.0, 3.0, 4.0, 1.5 waves
constexpr int kGeometryModels = 6;
double occupancy_threshold(int model);
// K for a shorter sequence of `cols` residues.
int k_for_cols(int cols, int alpha, int model);
// Bucket ids: position of K in kernel_k_list, plus once the list's length for
// pairs that span several strips with at most kMultiShortCols columns, twice for
// longer ones (bucket_ids() in all).
int bucket_ids();
// Host table shorter-length -> bucket id, 65536 entries, cached per (alpha, model).
const std::vector<uint8_t>& bucket_table(int alpha, int model);
// Key bits of the device radix sort: bucket id << 16 | (65535 - longer length).
constexpr int kSortBits = 22;
// Multi-strip pairs up to this many columns form buckets of their own (ids nk ..
// 2nk-1), longer ones the ids 2nk .. 3nk-1: 512 x 3 + 3 < 2048 keeps DNA buckets
// of the first class on the exact f16 kernels.
constexpr int kMultiShortCols = 512;
// Fill job.buckets from per-pair lengths (host), ordered by ascending K.
void count_buckets(Job& job, const int32_t* len_ref, const int32_t* len_qry);
// Workgroups of a full-occupancy batched launch for K columns per lane (the grid
// of a bucket is min(this, ceil(pairs / kPairsPerWave))).
int batched_grid_cap(const Exec& ex, int K, int alpha);
// Kernel of a batched bucket: the f16 kernel when its score bound allows, else
// the int16 one.
int batched_kind(const Bucket& bk, int32_t max_sub, bool allow_f16);
// Scratch bytes of the radix sort for n pairs.
int sort_scratch_bytes(int64_t n, size_t* bytes);

// A protein table with M[r][c] != M[c][r] for some pair (the reverse pass may
// then need its own maximum, kernel.cpp:444-459, 1021-1030).
bool table_is_asymmetric(const gbsw_scoring* sc, const Job& job);
// Validate a scoring for a job and build its substitution table; *p8: the job
// can use the int8 profile.
int prepare_scoring(const Job& job, const gbsw_scoring* sc, std::vector<int16_t>& sub, int* max_sub, bool* p8,
                    bool* f16p8 = nullptr);

// Allocation-free entry point (gbsw_align_noalloc): the device scratch is one
// static array per device (gbsw_noalloc.hip) of kNoallocScratchBytes.
constexpr size_t kNoallocScratchBytes = (size_t)256 << 20;
constexpr int kNoallocLaunchEvents = 128; // event pairs pre-created for its launches
// Set-up, once per device with that device current: the device address of the
// device's scratch array (hipGetSymbolAddress).  Thread-safe.
int noalloc_setup(int device);
// That address (nullptr before noalloc_setup succeeded on the device).
void* noalloc_scratch(int device);

// Enqueue a whole alignment job (binning if batched, kernels) on ex.stream.
// Asynchronous: ex.ev_stop is recorded after the last kernel and the status words
// are copied to ex.status_host.
int enqueue_job(Exec& ex, const Job& job, const gbsw_scoring* sc, int mode);
// After ex.ev_stop completed: recompute int16-saturated pairs with the int32
// kernel (synchronously, rare), check the status words and collect timings.
int complete_job(Exec& ex, const Job& job, const gbsw_scoring* sc, double* device_ms);

} // namespace gbsw
