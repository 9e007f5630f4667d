// gbsw_api.cpp -- C ABI of the engine (include/gbsw.h): contexts, resident
// batches, length binning and kernel launches.
//
// Replaces the host pipeline of driver.cpp:15-243 / 245-489 (packing, chunking,
// two-stream copies, host newMin between the passes, pinned result arrays) with:
// a batch uploaded once into HBM, a forward->reverse hand-off that never leaves
// the device, and status codes instead of exit().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gbsw.h"
#include "gbsw_kernels.hpp"

using namespace gbsw;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(GBSW_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

// Residue order of the reference's protein table (driver.cpp:252-258 maps each
// of these letters to its index; every other byte below 91 maps to 0).
const char kAaOrder[] = "ARNDCQEGHILKMFPSTWYVBZX*";

struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    ~DeviceBuffer() { release(); }
    void release()
    {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    hipError_t reserve(size_t b)
    {
        if (b <= bytes) return hipSuccess;
        release();
        hipError_t e = hipMalloc(&ptr, b);
        if (e == hipSuccess) bytes = b;
        else ptr = nullptr;
        return e;
    }
    template <class T> T* as() const { return (T*)ptr; }
};

// One length bucket of the batched form: a kernel geometry and the slice of the
// sorted order array it processes.
struct Bucket {
    int K = 0;
    int64_t begin = 0, count = 0; // slots in the order array
    int64_t cells = 0;
};

} // namespace

constexpr int kAuxStreams = 3;

// Streams used for the length buckets of the batched form: 1 + kAuxStreams by
// default; GBSW_STREAMS=1 serialises the buckets (per-kernel timing and counters).
int bucket_streams()
{
    static const int n = [] {
        const char* e = std::getenv("GBSW_STREAMS");
        const int v = e ? std::atoi(e) : kAuxStreams + 1;
        return std::max(1, std::min(kAuxStreams + 1, v));
    }();
    return n;
}

struct gbsw_context {
    int device = 0;
    int num_cu = 0;
    hipStream_t stream = nullptr;
    hipStream_t aux[kAuxStreams] = {nullptr, nullptr, nullptr}; // concurrent bucket launches
    hipEvent_t ev_fork = nullptr, ev_join[kAuxStreams] = {nullptr, nullptr, nullptr};
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<hipEvent_t> kev; // per-launch event pairs
    DeviceBuffer bnd, queue, status, sub, lut, colmax, redo;
};

struct gbsw_batch {
    gbsw_context* ctx = nullptr;
    int32_t alphabet = 0;
    int64_t n = 0;
    DeviceBuffer ref, ref_off, qry, qry_off, out, order;
    std::vector<int32_t> len_ref, len_qry;
    int32_t alpha = 0;           // residue codes in use
    uint8_t lut[256];            // byte -> code, 0xFF outside the alphabet
    int32_t max_rows = 0;        // longest sequence
    int32_t max_cols = 0;        // longest shorter-sequence of a pair
    int64_t cells = 0;
    // batched form: length buckets.  Counts come from the lengths at upload;
    // the work order itself is sorted on the device by every gbsw_batch_align.
    std::vector<Bucket> buckets;
    DeviceBuffer bucket_of_len, keys, keys_sorted, vals, sort_temp;
    size_t sort_temp_bytes = 0;
    gbsw_stats stats{};
    struct Launch {
        int kind, K, grid;
        int64_t pairs, cells;
        double ms;
    };
    std::vector<Launch> launches; // of the last gbsw_batch_align
};

namespace {

// LDS bytes of a workgroup; bnd_bytes = 4 (packed int16) or 8 (int32 kernel).
int lds_for(int K, int alpha, int bnd_bytes = 4)
{
    return lds_layout(kVirtPerPair * kstride(K), alpha, bnd_bytes).total;
}

// Largest compiled K whose LDS footprint fits a 64 KiB workgroup.
int max_k_for(int alpha)
{
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    int best = ks[0];
    for (int i = 0; i < nk; ++i)
        if (lds_for(ks[i], alpha) <= 65536) best = ks[i];
    return best;
}

// Geometry for a shorter sequence of `cols` residues: the K (columns per
// virtual lane) minimising strips x (K + 2), i.e. padded columns plus the fixed
// per-step cost of a step (about two columns' worth of work); ties go to fewer
// strips.
int k_for_cols(int cols, int kmax)
{
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    int best = ks[0];
    int64_t best_cost = INT64_MAX;
    for (int i = 0; i < nk; ++i) {
        const int K = ks[i];
        if (K > kmax) break;
        const int64_t strips = std::max<int64_t>(1, (cols + kVirtPerPair * K - 1) / (kVirtPerPair * K));
        const int64_t cost = strips * (K + 2) * 4 + (strips - 1);
        if (cost < best_cost) { best_cost = cost; best = K; }
    }
    return best;
}

int grid_for(gbsw_context* ctx, int kind, int K, int lds, int64_t nquads)
{
    int occ = kernel_occupancy(kind, K, lds);
    if (occ < 1) occ = 1;
    const int64_t full = (int64_t)occ * ctx->num_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(full, nquads));
}

// Length buckets of a batch: pairs whose shorter sequence gets the same K.
// Bucket i covers slots [begin, begin + count) of the device-sorted order.
int setup_buckets(gbsw_batch* b)
{
    const int kmax = max_k_for(b->alpha);
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    std::vector<uint8_t> bucket_of_len(65536);
    for (int m = 0; m < 65536; ++m) {
        const int K = k_for_cols(m, kmax);
        int idx = 0;
        while (idx < nk && ks[idx] != K) ++idx;
        bucket_of_len[(size_t)m] = (uint8_t)idx;
    }
    std::vector<int64_t> count((size_t)nk, 0), cells((size_t)nk, 0);
    for (int64_t p = 0; p < b->n; ++p) {
        const int32_t la = b->len_ref[(size_t)p], lb = b->len_qry[(size_t)p];
        const int bi = bucket_of_len[(size_t)std::min<int32_t>(std::min(la, lb), 65535)];
        count[(size_t)bi] += 1;
        cells[(size_t)bi] += (int64_t)la * lb;
    }
    b->buckets.clear();
    int64_t at = 0;
    for (int i = 0; i < nk; ++i) {
        if (count[(size_t)i] == 0) continue;
        Bucket bk;
        bk.K = ks[i];
        bk.begin = at;
        bk.count = count[(size_t)i];
        bk.cells = cells[(size_t)i];
        at += bk.count;
        b->buckets.push_back(bk);
    }
    const size_t n = (size_t)std::max<int64_t>(1, b->n);
    HIP_TRY(b->bucket_of_len.reserve(65536));
    HIP_TRY(hipMemcpy(b->bucket_of_len.ptr, bucket_of_len.data(), 65536, hipMemcpyHostToDevice));
    HIP_TRY(b->keys.reserve(n * 4));
    HIP_TRY(b->keys_sorted.reserve(n * 4));
    HIP_TRY(b->vals.reserve(n * 4));
    HIP_TRY(b->order.reserve(n * 4));
    size_t tb = 0;
    HIP_TRY(sort_pairs(nullptr, tb, nullptr, nullptr, nullptr, nullptr, (int)n, 20, nullptr));
    HIP_TRY(b->sort_temp.reserve(std::max<size_t>(tb, 16)));
    b->sort_temp_bytes = tb;
    return GBSW_OK;
}

// Device work order of the batched form: keys, then a stable radix sort.
int sort_work(gbsw_batch* b, hipStream_t stream)
{
    if (b->n <= 0) return GBSW_OK;
    HIP_TRY(bin_keys(b->ref_off.as<uint64_t>(), b->qry_off.as<uint64_t>(), b->n, b->bucket_of_len.as<uint8_t>(),
                     b->keys.as<uint32_t>(), b->vals.as<int32_t>(), stream));
    size_t tb = b->sort_temp_bytes;
    HIP_TRY(sort_pairs(b->sort_temp.ptr, tb, b->keys.as<uint32_t>(), b->keys_sorted.as<uint32_t>(), b->vals.as<int32_t>(),
                       b->order.as<int32_t>(), (int)b->n, 20, stream));
    return GBSW_OK;
}

int validate_scoring(const gbsw_batch* b, const gbsw_scoring* sc, std::vector<int16_t>& sub, int* max_sub)
{
    if (sc->alphabet != b->alphabet) return fail(GBSW_E_ARGUMENT, "scoring alphabet differs from the batch alphabet");
    if (sc->gap_open > 0 || sc->gap_extend > 0)
        return fail(GBSW_E_SCORING, "gap_open and gap_extend must be <= 0");
    const int A = b->alpha;
    sub.assign((size_t)std::max(1, A * A), 0);
    int mx = 0;
    if (b->alphabet == GBSW_DNA) {
        for (int r = 0; r < A; ++r)
            for (int c = 0; c < A; ++c) sub[(size_t)(r * A + c)] = r == c ? sc->match : sc->mismatch;
        mx = std::max<int>(sc->match, sc->mismatch);
    } else {
        for (int k = 0; k < 576; ++k) {
            sub[(size_t)k] = sc->matrix[k];
            mx = std::max<int>(mx, sc->matrix[k]);
        }
    }
    for (int16_t v : sub)
        if (v < -16000 || v > 16000) return fail(GBSW_E_SCORING, "substitution scores must lie in [-16000, 16000]");
    *max_sub = std::max(0, mx);
    return GBSW_OK;
}

KernelArgs base_args(gbsw_batch* b, const gbsw_scoring* sc, int max_sub)
{
    gbsw_context* ctx = b->ctx;
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ref = b->ref.as<uint8_t>();
    a.ref_off = b->ref_off.as<uint64_t>();
    a.qry = b->qry.as<uint8_t>();
    a.qry_off = b->qry_off.as<uint64_t>();
    a.n = b->n;
    a.out = b->out.as<int32_t>();
    a.status = ctx->status.as<int32_t>();
    a.redo = ctx->redo.as<int32_t>();
    a.lut = ctx->lut.as<uint8_t>();
    a.sub = ctx->sub.as<int16_t>();
    a.colmax = ctx->colmax.as<int16_t>();
    a.alpha = b->alpha;
    a.gap_min = std::min(-(int32_t)sc->gap_open, -(int32_t)sc->gap_extend);
    a.gap_open = sc->gap_open;
    a.gap_extend = sc->gap_extend;
    a.sat_limit = 32767 - max_sub;
    a.bnd_rows = std::max(1, b->max_rows);
    return a;
}

// Launch one kernel between two timing events.  `bnd` must hold
// grid * kPairsPerWave * bnd_rows words for this launch alone.
int launch_timed(gbsw_context* ctx, hipStream_t stream, int kind, int K, int grid, int lds, KernelArgs& a,
                 void* bnd, size_t ev_slot)
{
    a.bnd = bnd;
    while (ctx->kev.size() < 2 * (ev_slot + 1)) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        ctx->kev.push_back(e);
    }
    HIP_TRY(hipEventRecord(ctx->kev[2 * ev_slot], stream));
    HIP_TRY(kernel_launch(kind, K, grid, lds, stream, a));
    HIP_TRY(hipEventRecord(ctx->kev[2 * ev_slot + 1], stream));
    return GBSW_OK;
}

} // namespace

extern "C" {

const char* gbsw_last_error(void) { return g_err.c_str(); }

const char* gbsw_version(void) { return "gbsw-mi355x 0.1 (gfx950, packed int16 wavefront)"; }

int gbsw_device_memory(int32_t device, uint64_t* bytes)
{
    if (!bytes) return fail(GBSW_E_ARGUMENT, "bytes is null");
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    *bytes = (uint64_t)prop.totalGlobalMem;
    return GBSW_OK;
}

int gbsw_context_create(int32_t device, gbsw_context** out)
{
    if (!out) return fail(GBSW_E_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(GBSW_E_ARGUMENT, "no such device");
    HIP_TRY(hipSetDevice(device));
    gbsw_context* ctx = new gbsw_context();
    ctx->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_start);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_stop);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
    for (int i = 0; i < kAuxStreams && e == hipSuccess; ++i) {
        e = hipStreamCreateWithFlags(&ctx->aux[i], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join[i], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = ctx->status.reserve(sizeof(int32_t) * kStatusWords);
    if (e == hipSuccess) e = ctx->queue.reserve(sizeof(uint32_t) * 64);
    if (e == hipSuccess) e = ctx->sub.reserve(sizeof(int16_t) * 1024);
    if (e == hipSuccess) e = ctx->lut.reserve(256);
    if (e == hipSuccess) e = ctx->colmax.reserve(sizeof(int16_t) * 32);
    if (e != hipSuccess) {
        gbsw_context_destroy(ctx);
        return fail(GBSW_E_DEVICE, std::string("context setup: ") + hipGetErrorString(e));
    }
    ctx->num_cu = prop.multiProcessorCount;
    *out = ctx;
    return GBSW_OK;
}

int gbsw_context_destroy(gbsw_context* ctx)
{
    if (!ctx) return GBSW_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (hipEvent_t e : ctx->kev) (void)hipEventDestroy(e);
    if (ctx->ev_start) (void)hipEventDestroy(ctx->ev_start);
    if (ctx->ev_stop) (void)hipEventDestroy(ctx->ev_stop);
    if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
    for (int i = 0; i < kAuxStreams; ++i) {
        if (ctx->aux[i]) { (void)hipStreamSynchronize(ctx->aux[i]); (void)hipStreamDestroy(ctx->aux[i]); }
        if (ctx->ev_join[i]) (void)hipEventDestroy(ctx->ev_join[i]);
    }
    ctx->bnd.release(); ctx->queue.release(); ctx->status.release(); ctx->sub.release(); ctx->lut.release();
    ctx->colmax.release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return GBSW_OK;
}

int gbsw_batch_create(gbsw_context* ctx, int32_t alphabet, int64_t n, const char* ref,
                      const uint64_t* ref_off, const char* query, const uint64_t* query_off,
                      gbsw_batch** out)
{
    if (!ctx || !out || n < 0 || (n > 0 && (!ref_off || !query_off)))
        return fail(GBSW_E_ARGUMENT, "null argument or negative pair count");
    if (alphabet != GBSW_DNA && alphabet != GBSW_PROTEIN) return fail(GBSW_E_ARGUMENT, "unknown alphabet");
    if (n > (int64_t)0x7FFFFFFF) return fail(GBSW_E_RANGE, "more than 2^31-1 pairs in one batch");
    *out = nullptr;
    const uint64_t ref_bytes = n ? ref_off[n] - ref_off[0] : 0;
    const uint64_t qry_bytes = n ? query_off[n] - query_off[0] : 0;
    if ((ref_bytes && !ref) || (qry_bytes && !query)) return fail(GBSW_E_ARGUMENT, "sequence buffer is null");
    HIP_TRY(hipSetDevice(ctx->device));

    gbsw_batch* b = new gbsw_batch();
    b->ctx = ctx;
    b->alphabet = alphabet;
    b->n = n;
    b->len_ref.resize((size_t)n);
    b->len_qry.resize((size_t)n);
    std::vector<uint64_t> ro((size_t)n + 1), qo((size_t)n + 1);
    for (int64_t p = 0; p <= n; ++p) {
        ro[(size_t)p] = n ? ref_off[p] - ref_off[0] : 0;
        qo[(size_t)p] = n ? query_off[p] - query_off[0] : 0;
    }
    for (int64_t p = 0; p < n; ++p) {
        if (ro[(size_t)p + 1] < ro[(size_t)p] || qo[(size_t)p + 1] < qo[(size_t)p]) {
            delete b;
            return fail(GBSW_E_ARGUMENT, "offsets must be non-decreasing");
        }
        const uint64_t lr = ro[(size_t)p + 1] - ro[(size_t)p], lq = qo[(size_t)p + 1] - qo[(size_t)p];
        if (lr > 65000 || lq > 65000) {
            delete b;
            return fail(GBSW_E_RANGE, "sequences longer than 65000 residues are not supported");
        }
        b->len_ref[(size_t)p] = (int32_t)lr;
        b->len_qry[(size_t)p] = (int32_t)lq;
        b->max_rows = std::max<int32_t>(b->max_rows, (int32_t)std::max(lr, lq));
        b->max_cols = std::max<int32_t>(b->max_cols, (int32_t)std::min(lr, lq));
        b->cells += (int64_t)(lr * lq);
    }
    // residue alphabet: protein uses the reference encoding, DNA the set of bytes present
    std::memset(b->lut, 0xFF, sizeof(b->lut));
    const uint8_t* rp = (const uint8_t*)ref + (n ? ref_off[0] : 0);
    const uint8_t* qp = (const uint8_t*)query + (n ? query_off[0] : 0);
    if (alphabet == GBSW_PROTEIN) {
        for (int c = 0; c < 91; ++c) {
            const char* f = std::strchr(kAaOrder, c);
            b->lut[c] = (uint8_t)((f && c) ? (f - kAaOrder) : 0);
        }
        b->alpha = 24;
        bool bad = false;
        for (uint64_t i = 0; i < ref_bytes && !bad; ++i) bad = rp[i] >= 91;
        for (uint64_t i = 0; i < qry_bytes && !bad; ++i) bad = qp[i] >= 91;
        if (bad) {
            delete b;
            return fail(GBSW_E_RESIDUE, "protein residue byte >= 91 (the reference encoding table has 91 entries)");
        }
    } else {
        bool seen[256] = {false};
        for (uint64_t i = 0; i < ref_bytes; ++i) seen[rp[i]] = true;
        for (uint64_t i = 0; i < qry_bytes; ++i) seen[qp[i]] = true;
        int A = 0;
        for (int c = 0; c < 256; ++c)
            if (seen[c]) b->lut[c] = (uint8_t)A++;
        if (A > kMaxAlpha) {
            delete b;
            return fail(GBSW_E_RESIDUE, "more than 31 distinct DNA bytes in one batch");
        }
        b->alpha = std::max(A, 1);
    }
    hipError_t e = hipSuccess;
    e = b->ref.reserve(std::max<uint64_t>(ref_bytes, 16));
    if (e == hipSuccess) e = b->qry.reserve(std::max<uint64_t>(qry_bytes, 16));
    if (e == hipSuccess) e = b->ref_off.reserve(sizeof(uint64_t) * (size_t)(n + 1));
    if (e == hipSuccess) e = b->qry_off.reserve(sizeof(uint64_t) * (size_t)(n + 1));
    if (e == hipSuccess) e = b->out.reserve(sizeof(int32_t) * 5 * (size_t)std::max<int64_t>(n, 1));
    if (e == hipSuccess && ref_bytes) e = hipMemcpy(b->ref.ptr, rp, ref_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && qry_bytes) e = hipMemcpy(b->qry.ptr, qp, qry_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->ref_off.ptr, ro.data(), sizeof(uint64_t) * (size_t)(n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->qry_off.ptr, qo.data(), sizeof(uint64_t) * (size_t)(n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->out.ptr, 0, sizeof(int32_t) * 5 * (size_t)std::max<int64_t>(n, 1));
    if (e != hipSuccess) {
        delete b;
        return fail(e == hipErrorOutOfMemory ? GBSW_E_MEMORY : GBSW_E_DEVICE,
                    std::string("batch upload: ") + hipGetErrorString(e));
    }
    b->stats.cells = b->cells;
    const int brc = setup_buckets(b);
    if (brc) {
        const std::string err = g_err;
        delete b;
        g_err = err;
        return brc;
    }
    *out = b;
    return GBSW_OK;
}

int gbsw_batch_destroy(gbsw_batch* b)
{
    if (!b) return GBSW_OK;
    (void)hipSetDevice(b->ctx->device);
    delete b;
    return GBSW_OK;
}

int gbsw_batch_align(gbsw_batch* b, const gbsw_scoring* sc, int32_t mode)
{
    if (!b || !sc) return fail(GBSW_E_ARGUMENT, "null argument");
    if (mode != GBSW_PLAIN && mode != GBSW_BATCHED) return fail(GBSW_E_ARGUMENT, "unknown mode");
    gbsw_context* ctx = b->ctx;
    HIP_TRY(hipSetDevice(ctx->device));
    std::vector<int16_t> sub;
    int max_sub = 0;
    int rc = validate_scoring(b, sc, sub, &max_sub);
    if (rc) return rc;
    // result keys hold scores in 23 bits; int16 saturation is handled per pair
    if ((int64_t)b->max_cols * max_sub >= (1 << 23))
        return fail(GBSW_E_RANGE, "possible score exceeds 2^23");
    b->stats = gbsw_stats{};
    b->stats.cells = b->cells;
    if (b->n == 0) return GBSW_OK;

    HIP_TRY(hipMemcpyAsync(ctx->sub.ptr, sub.data(), sizeof(int16_t) * sub.size(), hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipMemcpyAsync(ctx->lut.ptr, b->lut, 256, hipMemcpyHostToDevice, ctx->stream));
    std::vector<int16_t> colmax((size_t)b->alpha, 0);
    for (int c = 0; c < b->alpha; ++c)
        for (int r = 0; r < b->alpha; ++r)
            colmax[(size_t)c] = std::max<int16_t>(colmax[(size_t)c], sub[(size_t)(r * b->alpha + c)]);
    HIP_TRY(hipMemcpyAsync(ctx->colmax.ptr, colmax.data(), sizeof(int16_t) * colmax.size(), hipMemcpyHostToDevice,
                           ctx->stream));
    HIP_TRY(ctx->redo.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(1, b->n)));
    KernelArgs a = base_args(b, sc, max_sub);
    const int kmax = max_k_for(b->alpha);
    std::vector<std::pair<int, int>> timed; // (event slot, is main kernel)
    HIP_TRY(zero_u32((uint32_t*)ctx->status.ptr, kStatusWords, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->ev_start, ctx->stream));
    size_t slot = 0;
    b->launches.clear();
    if (mode == GBSW_PLAIN) {
        // one geometry for the whole batch, sized by its longest shorter sequence
        const int K = k_for_cols(b->max_cols, kmax);
        const int lds = lds_for(K, b->alpha);
        const int64_t nquads = (b->n + kPairsPerWave - 1) / kPairsPerWave;
        a.order = nullptr;
        a.order_n = b->n;
        const int gf = grid_for(ctx, kPlainForward, K, lds, nquads);
        const int gr = grid_for(ctx, kPlainReverse, K, lds, nquads);
        HIP_TRY(ctx->bnd.reserve((size_t)std::max(gf, gr) * kPairsPerWave * (size_t)a.bnd_rows * sizeof(uint32_t)));
        rc = launch_timed(ctx, ctx->stream, kPlainForward, K, gf, lds, a, ctx->bnd.ptr, slot++);
        if (rc) return rc;
        b->launches.push_back({kPlainForward, K, gf, b->n, b->cells, 0.0});
        rc = launch_timed(ctx, ctx->stream, kPlainReverse, K, gr, lds, a, ctx->bnd.ptr, slot++);
        if (rc) return rc;
        b->launches.push_back({kPlainReverse, K, gr, b->n, 0, 0.0});
        b->stats.buckets = 1;
    } else {
        rc = sort_work(b, ctx->stream);
        if (rc) return rc;
        if (ctx->queue.bytes < sizeof(uint32_t) * b->buckets.size())
            HIP_TRY(ctx->queue.reserve(sizeof(uint32_t) * b->buckets.size()));
        HIP_TRY(zero_u32((uint32_t*)ctx->queue.ptr, (int)b->buckets.size(), ctx->stream));
        // Buckets run concurrently on 1 + kAuxStreams streams, the ones whose single
        // work item takes longest (largest K) first, so a bucket with few long
        // pairs overlaps the bulk instead of running alone at the end.
        std::vector<size_t> order(b->buckets.size());
        std::vector<int> grids(b->buckets.size()), ldss(b->buckets.size());
        size_t bnd_words = 0;
        for (size_t i = 0; i < b->buckets.size(); ++i) {
            order[i] = i;
            const Bucket& bk = b->buckets[i];
            ldss[i] = lds_for(bk.K, b->alpha);
            grids[i] = grid_for(ctx, kBatched, bk.K, ldss[i], (bk.count + kPairsPerWave - 1) / kPairsPerWave);
            bnd_words += (size_t)grids[i] * kPairsPerWave * (size_t)a.bnd_rows;
        }
        std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return b->buckets[x].K > b->buckets[y].K; });
        HIP_TRY(ctx->bnd.reserve(bnd_words * sizeof(uint32_t)));
        HIP_TRY(hipEventRecord(ctx->ev_fork, ctx->stream));
        for (int i = 0; i < kAuxStreams; ++i) HIP_TRY(hipStreamWaitEvent(ctx->aux[i], ctx->ev_fork, 0));
        size_t bnd_at = 0;
        for (size_t j = 0; j < order.size(); ++j) {
            const size_t i = order[j];
            const Bucket& bk = b->buckets[i];
            a.order = b->order.as<int32_t>() + bk.begin;
            a.order_n = bk.count;
            a.queue = (uint32_t*)ctx->queue.ptr + i;
            const int ns = bucket_streams();
            hipStream_t st = (j % ns) == 0 ? ctx->stream : ctx->aux[j % ns - 1];
            rc = launch_timed(ctx, st, kBatched, bk.K, grids[i], ldss[i], a, ctx->bnd.as<uint32_t>() + bnd_at, slot++);
            if (rc) return rc;
            bnd_at += (size_t)grids[i] * kPairsPerWave * (size_t)a.bnd_rows;
            b->launches.push_back({kBatched, bk.K, grids[i], bk.count, bk.cells, 0.0});
        }
        for (int i = 0; i < kAuxStreams; ++i) {
            HIP_TRY(hipEventRecord(ctx->ev_join[i], ctx->aux[i]));
            HIP_TRY(hipStreamWaitEvent(ctx->stream, ctx->ev_join[i], 0));
        }
        b->stats.buckets = (int32_t)b->buckets.size();
    }
    HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int32_t status[kStatusWords];
    HIP_TRY(hipMemcpy(status, ctx->status.ptr, sizeof(status), hipMemcpyDeviceToHost));
    if (status[kStatusSaturated] > 0) {
        // pairs whose int16 score reached the saturation limit: recompute them
        // (forward + reverse) with int32 cells
        int nk = 0;
        const int* wk = wide_k_list(&nk);
        int K = wk[nk - 1];
        for (int i = 0; i < nk; ++i)
            if (kVirtPerPair * wk[i] >= b->max_cols) { K = wk[i]; break; }
        const int lds = lds_for(K, b->alpha, 8);
        const int64_t count = status[kStatusSaturated];
        a.order = ctx->redo.as<int32_t>();
        a.order_n = count;
        a.sat_limit = INT_MAX;
        a.queue = (uint32_t*)ctx->queue.ptr;
        HIP_TRY(zero_u32(a.queue, 1, ctx->stream));
        const int grid = grid_for(ctx, kBatchedWide, K, lds, (count + kPairsPerWave - 1) / kPairsPerWave);
        HIP_TRY(ctx->bnd.reserve((size_t)grid * kPairsPerWave * (size_t)a.bnd_rows * 8));
        rc = launch_timed(ctx, ctx->stream, kBatchedWide, K, grid, lds, a, ctx->bnd.ptr, slot++);
        if (rc) return rc;
        b->launches.push_back({kBatchedWide, K, grid, count, 0, 0.0});
        HIP_TRY(hipEventRecord(ctx->ev_stop, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        HIP_TRY(hipMemcpy(status, ctx->status.ptr, sizeof(status), hipMemcpyDeviceToHost));
    }
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ctx->ev_start, ctx->ev_stop));
    b->stats.device_ms = ms;
    b->stats.launches = (int32_t)slot;
    double kmax_ms = 0.0;
    for (size_t i = 0; i < slot; ++i) {
        float t = 0.f;
        HIP_TRY(hipEventElapsedTime(&t, ctx->kev[2 * i], ctx->kev[2 * i + 1]));
        b->launches[i].ms = t;
        kmax_ms = std::max(kmax_ms, (double)t);
    }
    b->stats.main_kernel_ms = kmax_ms;
    if (status[kStatusRevMiss]) return fail(GBSW_E_INTERNAL, "reverse pass did not reach the forward score");
    return GBSW_OK;
}

int gbsw_batch_results(gbsw_batch* b, int32_t* score, int32_t* ref_begin, int32_t* ref_end,
                       int32_t* query_begin, int32_t* query_end)
{
    if (!b) return fail(GBSW_E_ARGUMENT, "batch is null");
    HIP_TRY(hipSetDevice(b->ctx->device));
    int32_t* dst[5] = {score, ref_begin, ref_end, query_begin, query_end};
    for (int f = 0; f < 5; ++f)
        if (dst[f] && b->n)
            HIP_TRY(hipMemcpy(dst[f], b->out.as<int32_t>() + (size_t)f * (size_t)b->n, sizeof(int32_t) * (size_t)b->n,
                              hipMemcpyDeviceToHost));
    return GBSW_OK;
}

int gbsw_batch_device_results(gbsw_batch* b, const int32_t** d_out)
{
    if (!b || !d_out) return fail(GBSW_E_ARGUMENT, "null argument");
    *d_out = b->out.as<int32_t>();
    return GBSW_OK;
}

int gbsw_batch_stats(const gbsw_batch* b, gbsw_stats* out)
{
    if (!b || !out) return fail(GBSW_E_ARGUMENT, "null argument");
    *out = b->stats;
    return GBSW_OK;
}

int gbsw_batch_launch_stats(const gbsw_batch* b, int32_t index, gbsw_launch_stat* out)
{
    if (!b || !out) return fail(GBSW_E_ARGUMENT, "null argument");
    if (index < 0 || (size_t)index >= b->launches.size()) return fail(GBSW_E_ARGUMENT, "launch index out of range");
    const gbsw_batch::Launch& L = b->launches[(size_t)index];
    out->kind = L.kind;
    out->columns_per_lane = L.K;
    out->workgroups = L.grid;
    out->pairs = L.pairs;
    out->cells = L.cells;
    out->ms = L.ms;
    return GBSW_OK;
}

int gbsw_align(gbsw_context* ctx, const gbsw_scoring* sc, int32_t mode, int64_t n, const char* ref,
               const uint64_t* ref_off, const char* query, const uint64_t* query_off, int32_t* score,
               int32_t* ref_begin, int32_t* ref_end, int32_t* query_begin, int32_t* query_end)
{
    if (!sc) return fail(GBSW_E_ARGUMENT, "scoring is null");
    gbsw_batch* b = nullptr;
    int rc = gbsw_batch_create(ctx, sc->alphabet, n, ref, ref_off, query, query_off, &b);
    if (rc) return rc;
    rc = gbsw_batch_align(b, sc, mode);
    if (!rc) rc = gbsw_batch_results(b, score, ref_begin, ref_end, query_begin, query_end);
    const std::string err = g_err;
    gbsw_batch_destroy(b);
    g_err = err;
    return rc;
}

} // extern "C"
