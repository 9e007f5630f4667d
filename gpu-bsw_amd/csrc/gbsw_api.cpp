// gbsw_api.cpp -- C ABI of the engine (include/gbsw.h): contexts and resident
// batches over the pipeline of gbsw_exec.cpp.
//
// Replaces the host pipeline of driver.cpp:15-243 / 245-489 (packing, chunking,
// two-stream copies, host newMin between the passes, pinned result arrays) with a
// batch uploaded once into HBM, a forward->reverse hand-off that never leaves the
// device, and status codes instead of exit() (utils.hpp:5-21).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gbsw_internal.hpp"

using namespace gbsw;

namespace gbsw {
const char* last_error_text();
}

namespace {

// Residue order of the reference's protein table (driver.cpp:252-258 maps each
// of these letters to its index; every other byte below 91 maps to 0).
const char kAaOrder[] = "ARNDCQEGHILKMFPSTWYVBZX*";

} // namespace

struct gbsw_batch;
struct gbsw_context {
    int device = 0;
    int num_cu = 0;
    Exec ex;
    gbsw_batch* oneshot = nullptr; // gbsw_align's batch, reused across calls
    // gbsw_align_noalloc: its own streams and events, and a batch whose device
    // buffers are views of the device's static scratch (set up at creation)
    Exec nx;
    gbsw_batch* nbatch = nullptr;
    std::string nx_error;          // why the allocation-free set-up failed ("" if it did not)
};

struct gbsw_batch {
    gbsw_context* ctx = nullptr;
    Job job;
    DeviceBuffer ref, ref_off, qry, qry_off, out, order;
    DeviceBuffer bucket_of_len, keys, keys_sorted, vals, sort_temp;
    std::vector<int32_t> len_ref, len_qry;
    uint8_t lut[256];
    uint8_t code_of[24]; // protein: reference code of each residue code present
    gbsw_stats stats{};
    std::vector<LaunchRec> launches; // of the last gbsw_batch_align
    JobGraph graph;
    bool one_shot = false; // a gbsw_align batch: aligned once, never worth a graph
    bool aligned = false;  // results of a gbsw_batch_align are in `out`
    // alignment paths: per-pair tables, run buffers, tightness scratch
    DeviceBuffer tb_meta, tb_runs, tb_dir, tb_dst, tb_tables;
    PinnedBuffer tb_host;
    std::vector<hipEvent_t> tb_events;
    ~gbsw_batch()
    {
        for (hipEvent_t e : tb_events) (void)hipEventDestroy(e);
    }
};

namespace gbsw {

// Residue table of an alphabet: protein uses the reference encoding; DNA maps each
// distinct byte present in `bytes` to its own code (raw-byte equality).  Returns
// the number of codes, or a negative status.
int build_alphabet(int32_t alphabet, const uint8_t* const* bytes, const uint64_t* lens, int nbuf, uint8_t* lut)
{
    std::memset(lut, 0xFF, 256);
    if (alphabet == GBSW_PROTEIN) {
        for (int c = 0; c < 91; ++c) {
            const char* f = std::strchr(kAaOrder, c);
            lut[c] = (uint8_t)((f && c) ? (f - kAaOrder) : 0);
        }
        for (int k = 0; k < nbuf; ++k)
            for (uint64_t i = 0; i < lens[k]; ++i)
                if (bytes[k][i] >= 91)
                    return fail(GBSW_E_RESIDUE, "protein residue byte >= 91 (the reference encoding table has 91 entries)");
        return 24;
    }
    bool seen[256] = {false};
    for (int k = 0; k < nbuf; ++k)
        for (uint64_t i = 0; i < lens[k]; ++i) seen[bytes[k][i]] = true;
    int A = 0;
    for (int c = 0; c < 256; ++c)
        if (seen[c]) lut[c] = (uint8_t)A++;
    if (A > kMaxAlpha) return fail(GBSW_E_RESIDUE, "more than 31 distinct DNA bytes in one batch");
    return std::max(A, 1);
}

// build_alphabet from the set of byte values present (bit b of mask[b / 32]),
// as scanned on the device after the upload.  Protein: only the reference codes
// that occur get a residue code (in code order), so the kernels' profiles hold
// no rows for absent residues; code_of[k] is the reference code (row of the
// caller's 24x24 matrix) of residue code k.
int alphabet_from_mask(int32_t alphabet, const uint32_t* mask, uint8_t* lut, uint8_t* code_of)
{
    auto present = [&](int c) { return (mask[c >> 5] >> (c & 31)) & 1u; };
    std::memset(lut, 0xFF, 256);
    if (alphabet == GBSW_PROTEIN) {
        for (int c = 91; c < 256; ++c)
            if (present(c))
                return fail(GBSW_E_RESIDUE, "protein residue byte >= 91 (the reference encoding table has 91 entries)");
        uint8_t enc[256];
        build_alphabet(alphabet, nullptr, nullptr, 0, enc);
        bool used[24] = {false};
        for (int c = 0; c < 91; ++c)
            if (present(c)) used[enc[c]] = true;
        uint8_t compact[24];
        int A = 0;
        for (int k = 0; k < 24; ++k)
            if (used[k]) {
                compact[k] = (uint8_t)A;
                code_of[A++] = (uint8_t)k;
            }
        if (A == 0) code_of[A++] = 0; // an empty batch: one code
        for (int c = 0; c < 91; ++c) lut[c] = used[enc[c]] ? compact[enc[c]] : 0xFF;
        return A;
    }
    int A = 0;
    for (int c = 0; c < 256; ++c)
        if (present(c)) lut[c] = (uint8_t)A++;
    if (A > kMaxAlpha) return fail(GBSW_E_RESIDUE, "more than 31 distinct DNA bytes in one batch");
    return std::max(A, 1);
}

} // namespace gbsw

namespace gbsw {

// Upload pairs into batch b (new, or one reused: its device buffers only grow),
// with ex's stream and geometry model.
static int batch_fill(gbsw_batch* b, gbsw_context* ctx, Exec& ex, int32_t alphabet, int64_t n, const char* ref,
                      const uint64_t* ref_off, const char* query, const uint64_t* query_off)
{
    if (!ctx || n < 0 || (n > 0 && (!ref_off || !query_off)))
        return fail(GBSW_E_ARGUMENT, "null argument or negative pair count");
    if (alphabet != GBSW_DNA && alphabet != GBSW_PROTEIN) return fail(GBSW_E_ARGUMENT, "unknown alphabet");
    if (n > (int64_t)0x7FFFFFFF) return fail(GBSW_E_RANGE, "more than 2^31-1 pairs in one batch");
    const uint64_t ref_bytes = n ? ref_off[n] - ref_off[0] : 0;
    const uint64_t qry_bytes = n ? query_off[n] - query_off[0] : 0;
    if ((ref_bytes && !ref) || (qry_bytes && !query)) return fail(GBSW_E_ARGUMENT, "sequence buffer is null");
    GBSW_HIP_TRY(hipSetDevice(ctx->device));

    b->ctx = ctx;
    b->job = Job{};
    b->graph.reset();
    b->aligned = false;
    b->stats = gbsw_stats{};
    b->launches.clear();
    Job& job = b->job;
    job.alphabet = alphabet;
    job.n = n;
    b->len_ref.resize((size_t)n);
    b->len_qry.resize((size_t)n);
    std::vector<uint64_t> ro((size_t)n + 1), qo((size_t)n + 1);
    for (int64_t p = 0; p <= n; ++p) {
        ro[(size_t)p] = n ? ref_off[p] - ref_off[0] : 0;
        qo[(size_t)p] = n ? query_off[p] - query_off[0] : 0;
    }
    for (int64_t p = 0; p < n; ++p) {
        if (ro[(size_t)p + 1] < ro[(size_t)p] || qo[(size_t)p + 1] < qo[(size_t)p]) {
            return fail(GBSW_E_ARGUMENT, "offsets must be non-decreasing");
        }
        const uint64_t lr = ro[(size_t)p + 1] - ro[(size_t)p], lq = qo[(size_t)p + 1] - qo[(size_t)p];
        if (lr > 65000 || lq > 65000) {
            return fail(GBSW_E_RANGE, "sequences longer than 65000 residues are not supported");
        }
        b->len_ref[(size_t)p] = (int32_t)lr;
        b->len_qry[(size_t)p] = (int32_t)lq;
        job.max_rows = std::max<int32_t>(job.max_rows, (int32_t)std::max(lr, lq));
        job.max_cols = std::max<int32_t>(job.max_cols, (int32_t)std::min(lr, lq));
        job.cells += (int64_t)(lr * lq);
    }
    const uint8_t* rp = (const uint8_t*)ref + (n ? ref_off[0] : 0);
    const uint8_t* qp = (const uint8_t*)query + (n ? query_off[0] : 0);
    job.lut_host = b->lut;
    job.kmodel = ex.kmodel;
    size_t sort_bytes = 0;
    int rc = sort_scratch_bytes(n, &sort_bytes);
    if (rc) return rc;
    const size_t np = (size_t)std::max<int64_t>(n, 1);
    hipError_t e = b->ref.reserve(std::max<uint64_t>(ref_bytes, 16));
    if (e == hipSuccess) e = b->qry.reserve(std::max<uint64_t>(qry_bytes, 16));
    if (e == hipSuccess) e = b->ref_off.reserve(sizeof(uint64_t) * (np + 1));
    if (e == hipSuccess) e = b->qry_off.reserve(sizeof(uint64_t) * (np + 1));
    if (e == hipSuccess) e = b->out.reserve(sizeof(int32_t) * 5 * np);
    if (e == hipSuccess) e = b->keys.reserve(std::max<size_t>(4 * np, 32));
    if (e == hipSuccess) e = b->keys_sorted.reserve(4 * np);
    if (e == hipSuccess) e = b->vals.reserve(4 * np);
    if (e == hipSuccess) e = b->order.reserve(4 * np);
    if (e == hipSuccess) e = b->sort_temp.reserve(sort_bytes);
    if (e == hipSuccess) e = b->bucket_of_len.reserve(65536);
    if (e == hipSuccess && ref_bytes) e = hipMemcpy(b->ref.ptr, rp, ref_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess && qry_bytes) e = hipMemcpy(b->qry.ptr, qp, qry_bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->ref_off.ptr, ro.data(), sizeof(uint64_t) * (size_t)(n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(b->qry_off.ptr, qo.data(), sizeof(uint64_t) * (size_t)(n + 1), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(b->out.ptr, 0, sizeof(int32_t) * 5 * np);
    // The residue codes need the set of byte values in the batch (every distinct
    // DNA byte is a code; protein bytes must be < 91): a scan of the uploaded
    // bytes on the device instead of a pass over them on the host.
    uint32_t seen[8] = {0};
    if (e == hipSuccess) {
        hipStream_t st = ex.stream;
        uint32_t* dmask = b->keys.as<uint32_t>();
        e = zero_u32(dmask, 8, st);
        if (e == hipSuccess) e = byte_presence(b->ref.as<uint8_t>(), ref_bytes, dmask, st);
        if (e == hipSuccess) e = byte_presence(b->qry.as<uint8_t>(), qry_bytes, dmask, st);
        if (e == hipSuccess) e = hipMemcpyAsync(seen, dmask, sizeof(seen), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GBSW_E_MEMORY : GBSW_E_DEVICE,
                    std::string("batch upload: ") + hipGetErrorString(e));
    {
        const int A = alphabet_from_mask(alphabet, seen, b->lut, b->code_of);
        if (A < 0) return A;
        job.alpha = A;
        job.code_of = alphabet == GBSW_PROTEIN ? b->code_of : nullptr;
    }
    count_buckets(job, b->len_ref.data(), b->len_qry.data());
    e = hipMemcpy(b->bucket_of_len.ptr, bucket_table(job.alpha, job.kmodel).data(), 65536, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(GBSW_E_DEVICE, std::string("batch upload: ") + hipGetErrorString(e));
    job.ref = b->ref.as<uint8_t>();
    job.qry = b->qry.as<uint8_t>();
    job.ref_off = b->ref_off.as<uint64_t>();
    job.qry_off = b->qry_off.as<uint64_t>();
    job.out = b->out.as<int32_t>();
    job.bucket_of_len = b->bucket_of_len.as<uint8_t>();
    job.keys = b->keys.as<uint32_t>();
    job.keys_sorted = b->keys_sorted.as<uint32_t>();
    job.vals = b->vals.as<int32_t>();
    job.order = b->order.as<int32_t>();
    job.sort_temp = b->sort_temp.ptr;
    job.sort_temp_bytes = sort_bytes;
    b->stats.cells = job.cells;
    return GBSW_OK;
}

} // namespace gbsw

namespace {

// Both passes of every pair of batch b on ex's streams, blocking.
int batch_align_on(gbsw_batch* b, Exec& ex, const gbsw_scoring* sc, int32_t mode)
{
    if (mode != GBSW_PLAIN && mode != GBSW_BATCHED) return fail(GBSW_E_ARGUMENT, "unknown mode");
    GBSW_HIP_TRY(hipSetDevice(b->ctx->device));
    b->stats = gbsw_stats{};
    b->stats.cells = b->job.cells;
    b->launches.clear();
    // a resident batch captures its work into a graph; gbsw_align's one-off
    // batch does not
    b->job.graph = b->one_shot ? nullptr : &b->graph;
    int rc = enqueue_job(ex, b->job, sc, mode);
    if (rc) {
        (void)hipStreamSynchronize(ex.stream);
        return rc;
    }
    double ms = 0.0;
    rc = complete_job(ex, b->job, sc, &ms);
    b->launches = ex.launches;
    b->stats.device_ms = ms;
    b->stats.launches = (int32_t)b->launches.size();
    b->stats.buckets = mode == GBSW_PLAIN ? 1 : (int32_t)b->job.buckets.size();
    b->stats.graph = ex.graph_state;
    for (const LaunchRec& L : b->launches) b->stats.main_kernel_ms = std::max(b->stats.main_kernel_ms, L.ms);
    b->aligned = rc == GBSW_OK;
    return rc;
}

// Carve the device scratch of gbsw_align_noalloc for one call of n pairs:
// tables, then the batch's arrays, then the redo list; what is left holds the
// strip-boundary rows (7/8) and the split form's bookkeeping (1/8).  Every
// buffer becomes a view (never grown, never freed).
int noalloc_carve(gbsw_context* ctx, int64_t n, uint64_t ref_bytes, uint64_t qry_bytes)
{
    uint8_t* base = (uint8_t*)noalloc_scratch(ctx->device);
    if (!base || !ctx->nx.stream)
        return fail(GBSW_E_DEVICE, "the allocation-free scratch is not set up on this device: " + ctx->nx_error);
    size_t sort_bytes = 0;
    if (int rc = sort_scratch_bytes(n, &sort_bytes)) return rc;
    const size_t np = (size_t)std::max<int64_t>(n, 1);
    size_t at = 0;
    auto take = [&](DeviceBuffer& d, size_t bytes) {
        const size_t p = at;
        at = (at + bytes + 255) & ~(size_t)255;
        if (at <= kNoallocScratchBytes) d.view(base + p, bytes);
    };
    Exec& x = ctx->nx;
    gbsw_batch* b = ctx->nbatch;
    take(x.status, sizeof(int32_t) * kStatusWords);
    take(x.queue, sizeof(uint32_t) * kXcds * 64);
    take(x.sub, sizeof(int16_t) * 1024);
    take(x.lut, 256);
    take(x.colmax, sizeof(int16_t) * 32);
    take(b->bucket_of_len, 65536);
    take(b->ref, std::max<uint64_t>(ref_bytes, 16));
    take(b->qry, std::max<uint64_t>(qry_bytes, 16));
    take(b->ref_off, sizeof(uint64_t) * (np + 1));
    take(b->qry_off, sizeof(uint64_t) * (np + 1));
    take(b->out, sizeof(int32_t) * 5 * np);
    take(b->keys, std::max<size_t>(4 * np, 32));
    take(b->keys_sorted, 4 * np);
    take(b->vals, 4 * np);
    take(b->order, 4 * np);
    take(b->sort_temp, sort_bytes);
    take(x.redo, sizeof(int32_t) * np);
    // at least 1 MiB for boundary rows (a wide pair on one workgroup)
    constexpr size_t kMinRest = (size_t)1 << 20;
    if (at + kMinRest > kNoallocScratchBytes)
        return fail(GBSW_E_RANGE, "gbsw_align_noalloc: " + std::to_string(n) + " pairs of " +
                                      std::to_string(ref_bytes + qry_bytes) + " residue bytes need " +
                                      std::to_string(at + kMinRest) + " bytes of scratch; the limit is " +
                                      std::to_string(kNoallocScratchBytes));
    const size_t rest = kNoallocScratchBytes - at;
    const size_t split_bytes = (rest / 8) & ~(size_t)255;
    take(x.split_buf, split_bytes);
    take(x.bnd, kNoallocScratchBytes - at);
    return GBSW_OK;
}

} // namespace

extern "C" {

const char* gbsw_last_error(void) { return gbsw::last_error_text(); }

const char* gbsw_version(void) { return "gbsw-mi355x 0.2 (gfx950, packed int16 wavefront)"; }

int gbsw_device_memory(int32_t device, uint64_t* bytes)
{
    if (!bytes) return fail(GBSW_E_ARGUMENT, "bytes is null");
    hipDeviceProp_t prop;
    GBSW_HIP_TRY(hipGetDeviceProperties(&prop, device));
    *bytes = (uint64_t)prop.totalGlobalMem;
    return GBSW_OK;
}

int gbsw_context_create(int32_t device, gbsw_context** out)
{
    if (!out) return fail(GBSW_E_ARGUMENT, "out is null");
    *out = nullptr;
    int count = 0;
    GBSW_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(GBSW_E_ARGUMENT, "no such device");
    GBSW_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    GBSW_HIP_TRY(hipGetDeviceProperties(&prop, device));
    gbsw_context* ctx = new gbsw_context();
    ctx->device = device;
    ctx->num_cu = prop.multiProcessorCount;
    const int rc = ctx->ex.init(device, ctx->num_cu);
    if (rc) {
        ctx->ex.destroy();
        delete ctx;
        return rc;
    }
    // set-up of the allocation-free entry point: the device address of this
    // device's scratch, its streams and events; a failure here only disables
    // gbsw_align_noalloc (which then reports it)
    int nrc = noalloc_setup(device);
    if (!nrc) nrc = ctx->nx.init_views(device, ctx->num_cu);
    if (nrc) ctx->nx_error = gbsw_last_error();
    ctx->nbatch = new gbsw_batch();
    ctx->nbatch->one_shot = true;
    set_error("");
    *out = ctx;
    return GBSW_OK;
}

int gbsw_context_set_option(gbsw_context* ctx, int32_t option, int64_t value)
{
    if (!ctx) return fail(GBSW_E_ARGUMENT, "context is null");
    switch (option) {
    case GBSW_OPT_STREAMS:
        if (value < 1 || value > kAuxStreams + 1) return fail(GBSW_E_ARGUMENT, "streams must be 1..8");
        if (int rc = ctx->ex.ensure_streams((int)value)) return rc;
        ctx->ex.streams = (int)value;
        return GBSW_OK;
    case GBSW_OPT_F16:
        if (value != 0 && value != 1) return fail(GBSW_E_ARGUMENT, "f16 option is 0 or 1");
        ctx->ex.f16 = value != 0;
        return GBSW_OK;
    case GBSW_OPT_GRAPHS:
        if (value != 0 && value != 1) return fail(GBSW_E_ARGUMENT, "graphs option is 0 or 1");
        ctx->ex.graphs = value != 0;
        return GBSW_OK;
    case GBSW_OPT_SPLIT:
        if (value != 0 && value != 1) return fail(GBSW_E_ARGUMENT, "split option is 0 or 1");
        ctx->ex.split = value != 0;
        return GBSW_OK;
    case GBSW_OPT_GEOMETRY:
        if (value < 0 || value >= kGeometryModels) return fail(GBSW_E_ARGUMENT, "geometry model is 0..5");
        ctx->ex.kmodel = (int)value;
        return GBSW_OK;
    default:
        return fail(GBSW_E_ARGUMENT, "unknown option");
    }
}

int gbsw_context_destroy(gbsw_context* ctx)
{
    if (!ctx) return GBSW_OK;
    (void)hipSetDevice(ctx->device);
    delete ctx->oneshot;
    delete ctx->nbatch;
    ctx->nx.destroy();
    ctx->ex.destroy();
    delete ctx;
    return GBSW_OK;
}

int gbsw_batch_create(gbsw_context* ctx, int32_t alphabet, int64_t n, const char* ref,
                      const uint64_t* ref_off, const char* query, const uint64_t* query_off,
                      gbsw_batch** out)
{
    if (!out) return fail(GBSW_E_ARGUMENT, "out is null");
    *out = nullptr;
    gbsw_batch* b = new gbsw_batch();
    const int rc = ctx ? batch_fill(b, ctx, ctx->ex, alphabet, n, ref, ref_off, query, query_off)
                       : fail(GBSW_E_ARGUMENT, "context is null");
    if (rc) {
        const std::string err = gbsw_last_error();
        delete b;
        set_error(err);
        return rc;
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
    return batch_align_on(b, b->ctx->ex, sc, mode);
}

int gbsw_batch_traceback(gbsw_batch* b, const gbsw_scoring* sc, int64_t capacity, uint32_t* cigar,
                         int64_t* cigar_offsets, int32_t* path_score, double* device_ms)
{
    if (!b || !sc || !cigar_offsets) return fail(GBSW_E_ARGUMENT, "null argument");
    if (capacity < 0 || (capacity > 0 && !cigar)) return fail(GBSW_E_ARGUMENT, "bad cigar buffer");
    if (!b->aligned) return fail(GBSW_E_ARGUMENT, "the batch has no alignment results (gbsw_batch_align first)");
    if (sc->gap_open > sc->gap_extend)
        return fail(GBSW_E_SCORING, "alignment paths need gap_open <= gap_extend (a gap run is one gap)");
    gbsw_context* ctx = b->ctx;
    GBSW_HIP_TRY(hipSetDevice(ctx->device));
    const Job& job = b->job;
    std::vector<int16_t> sub;
    int32_t max_sub = 0;
    bool p8 = false;
    if (int rc = prepare_scoring(job, sc, sub, &max_sub, &p8)) return rc;
    const int64_t n = job.n;
    cigar_offsets[0] = 0;
    if (n == 0) {
        if (device_ms) *device_ms = 0.0;
        return GBSW_OK;
    }
    hipStream_t st = ctx->ex.stream;
    // host side in pinned memory, kept with the batch: results (5n int32), then
    // cap_off (n+1), order (n) and dir_off (n) int64, then run counts (n int32)
    const size_t hres = sizeof(int32_t) * (size_t)(5 * n), h64 = sizeof(int64_t) * (size_t)(3 * n + 1);
    GBSW_HIP_TRY(b->tb_host.reserve(hres + h64 + sizeof(int32_t) * (size_t)n));
    int32_t* res = b->tb_host.as<int32_t>();
    GBSW_HIP_TRY(hipMemcpyAsync(res, b->out.ptr, hres, hipMemcpyDeviceToHost, st));
    GBSW_HIP_TRY(hipStreamSynchronize(st));
    // cells are int32: |values| stay below (rows + columns) x the largest |score|
    int32_t big = std::max(std::abs(sc->gap_open), std::abs(sc->gap_extend));
    for (int16_t v : sub) big = std::max<int32_t>(big, std::abs((int32_t)v));
    // per pair: run capacity (rows + columns) and scratch region
    constexpr int64_t kTbScratch = (int64_t)4 << 30;
    int64_t* cap_off = (int64_t*)((uint8_t*)res + hres);
    int64_t* order = cap_off + n + 1;
    int64_t* dir_off = order + n;
    int32_t* nruns = (int32_t*)(dir_off + n);
    std::vector<int64_t> region((size_t)n);
    // class of a pair: its stripe count (capped at 127), then its rows in steps of 8
    // below 2048 and of 128 above (capped): the four pairs of a DP wave run the
    // largest stripe and row counts of the four, so like pairs go together
    constexpr int kTbClasses = 128 * 512;
    std::vector<uint16_t> cls((size_t)n);
    std::vector<int64_t> cls_count(kTbClasses, 0);
    cap_off[0] = 0;
    int64_t max_span = 0; // largest rows + columns of a rectangle
    for (int64_t p = 0; p < n; ++p) {
        const bool on = res[(size_t)p] > 0;
        const int32_t m = on ? res[(size_t)(2 * n + p)] - res[(size_t)(n + p)] : 0;
        const int32_t w = on ? res[(size_t)(4 * n + p)] - res[(size_t)(3 * n + p)] : 0;
        if (on && (m <= 0 || w <= 0)) return fail(GBSW_E_INTERNAL, "a pair with a positive score has an empty rectangle");
        if ((int64_t)(m + w) * big >= ((int64_t)1 << 28)) return fail(GBSW_E_RANGE, "path scores may exceed the int32 cells' range");
        max_span = std::max<int64_t>(max_span, (int64_t)m + w);
        cap_off[(size_t)p + 1] = cap_off[(size_t)p] + m + w;
        region[(size_t)p] = tb_region_bytes(m, w);
        const int32_t stripes = on ? std::min(127, (w + kTbStripeColumns - 1) / kTbStripeColumns) : 0;
        const int32_t rows = !on ? 0 : m < 2048 ? m >> 3 : std::min(511, 256 + ((m - 2048) >> 7));
        const int c = stripes * 512 + rows;
        cls[(size_t)p] = (uint16_t)c;
        ++cls_count[(size_t)c];
    }
    // order: largest classes first (their DP lasts longest; the short ones fill in
    // behind them), by a counting sort over the classes
    {
        std::vector<int64_t> start(kTbClasses);
        int64_t acc = 0;
        for (int c = kTbClasses - 1; c >= 0; --c) { start[(size_t)c] = acc; acc += cls_count[(size_t)c]; }
        for (int64_t p = 0; p < n; ++p) order[(size_t)start[cls[(size_t)p]]++] = p;
    }
    // int16 cells when every value stays inside +-32,000: a path value lies within
    // (rows + columns + 2) x the largest |score| of zero, and each sum formed from
    // one adds one score more
    const bool cells16 = (max_span + 4) * big < 32000;
    // launches in that order: one ends past kTbScratch bytes of scratch (scratch_soft)
    // once it holds enough waves to fill the GPU twice over (16 per SIMD, against
    // the DP's seven resident; the largest pairs come first, and a launch of few of
    // them runs latency-bound; 6, 8, 10, 12, 14 and 16 measured 1,225-1,231, 1,250-
    // 1,279, 1,271-1,304, 1,277-1,280, 1,301-1,308 and 1,304-1,310 GCUPS on the
    // 1M-pair DNA mix), at twice that many pairs, and in any case before
    // scratch_max (or at one pair): 12 GiB, or a third of the free device memory
    // when less (the scratch of two launches is held at once)
    size_t free_bytes = 0, total_bytes = 0;
    if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) {
        (void)hipGetLastError();
        free_bytes = (size_t)24 << 30;
    }
    const int64_t scratch_max =
        std::max<int64_t>((int64_t)256 << 20, std::min<int64_t>((int64_t)12 << 30, (int64_t)(free_bytes / 3)));
    const int64_t scratch_soft = std::min(kTbScratch, scratch_max);
    const int64_t fill_pairs = (int64_t)ctx->num_cu * 4 * 16 * (cells16 ? 8 : 4);
    std::vector<std::pair<int64_t, int64_t>> launches; // (first position, count)
    int64_t chunk_first = 0, chunk_bytes = 0, max_chunk = 0;
    for (int64_t x = 0; x < n; ++x) {
        const int64_t bytes = region[(size_t)order[(size_t)x]];
        const int64_t held = x - chunk_first;
        const bool full = chunk_bytes + bytes > scratch_max ||
                          (held >= fill_pairs && (chunk_bytes + bytes > scratch_soft || held >= 2 * fill_pairs));
        if (full && x > chunk_first) {
            launches.emplace_back(chunk_first, x - chunk_first);
            max_chunk = std::max(max_chunk, chunk_bytes);
            chunk_first = x;
            chunk_bytes = 0;
        }
        dir_off[(size_t)x] = chunk_bytes;
        chunk_bytes += bytes;
    }
    launches.emplace_back(chunk_first, n - chunk_first);
    max_chunk = std::max(max_chunk, chunk_bytes);
    // device tables: cap_off (n+1), dir_off (n), dst_off (n+1) int64; nruns, path
    // scores (n) int32; lut and substitution table
    const size_t meta_bytes = sizeof(int64_t) * (size_t)(4 * n + 2) + sizeof(int32_t) * (size_t)(2 * n);
    GBSW_HIP_TRY(b->tb_meta.reserve(meta_bytes));
    GBSW_HIP_TRY(b->tb_runs.reserve(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, cap_off[(size_t)n])));
    // two scratch halves: the walk of one launch overlaps the DP of the next
    const size_t half = ((size_t)std::max<int64_t>(8, max_chunk) + 255) & ~(size_t)255;
    GBSW_HIP_TRY(b->tb_dir.reserve(2 * half));
    GBSW_HIP_TRY(b->tb_tables.reserve(256 + sizeof(int16_t) * sub.size()));
    // same order as the host staging: cap_off, order, dir_off (one copy), then dst_off
    int64_t* d_cap = b->tb_meta.as<int64_t>();
    int64_t* d_order = d_cap + n + 1;
    int64_t* d_dir_off = d_order + n;
    int64_t* d_dst_off = d_dir_off + n;
    int32_t* d_nruns = (int32_t*)(d_dst_off + n + 1);
    int32_t* d_score = d_nruns + n;
    uint8_t* d_lut = b->tb_tables.as<uint8_t>();
    int16_t* d_sub = (int16_t*)(d_lut + 256);
    GBSW_HIP_TRY(hipMemcpyAsync(d_cap, cap_off, sizeof(int64_t) * (size_t)(3 * n + 1), hipMemcpyHostToDevice, st));
    GBSW_HIP_TRY(hipMemcpyAsync(d_lut, b->lut, 256, hipMemcpyHostToDevice, st));
    GBSW_HIP_TRY(hipMemcpyAsync(d_sub, sub.data(), sizeof(int16_t) * sub.size(), hipMemcpyHostToDevice, st));
    TbArgs a{};
    a.ref = b->ref.as<uint8_t>();
    a.qry = b->qry.as<uint8_t>();
    a.ref_off = b->ref_off.as<uint64_t>();
    a.qry_off = b->qry_off.as<uint64_t>();
    a.out = b->out.as<int32_t>();
    a.n = n;
    a.lut = d_lut;
    a.sub = d_sub;
    a.alpha = job.alpha;
    a.gap_open = sc->gap_open;
    a.gap_extend = sc->gap_extend;
    a.dir = b->tb_dir.as<uint8_t>();
    a.cap_off = d_cap;
    a.runs = b->tb_runs.as<uint32_t>();
    a.nruns = d_nruns;
    a.path_score = d_score;
    a.order = d_order;
    a.cells16 = cells16 ? 1 : 0;
    Exec& ex = ctx->ex;
    // Two pipelines, the launches at even and at odd positions, each with its own
    // scratch half, DP stream and walk stream (the context stream and three
    // auxiliary ones): launch c's walk waits for its DP, and launch c + 2's DP
    // (same half) for that walk.  A launch lasts as long as its longest wave (the
    // largest pairs come first), so the other pipeline's launches fill the GPU
    // while one runs out its tail.
    if (int rc = ex.ensure_streams(4)) return rc;
    GBSW_HIP_TRY(hipEventRecord(ex.ev_start, st));
    const hipStream_t dp_st[2] = {st, ex.aux[1]}, walk_st[2] = {ex.aux[0], ex.aux[2]};
    const size_t nev = 2 * launches.size() + 1; // DP c done, walk c done; the tables uploaded
    while (b->tb_events.size() < nev) {
        hipEvent_t e;
        GBSW_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        b->tb_events.push_back(e);
    }
    hipEvent_t uploaded = b->tb_events[nev - 1];
    GBSW_HIP_TRY(hipEventRecord(uploaded, st));
    for (int x = 0; x < 3; ++x) GBSW_HIP_TRY(hipStreamWaitEvent(ex.aux[x], uploaded, 0));
    for (size_t c = 0; c < launches.size(); ++c) {
        const size_t h = c & 1;
        a.p0 = launches[c].first;
        a.dir_off = d_dir_off + launches[c].first;
        a.dir = b->tb_dir.as<uint8_t>() + h * half;
        if (c >= 2) GBSW_HIP_TRY(hipStreamWaitEvent(dp_st[h], b->tb_events[2 * (c - 2) + 1], 0));
        GBSW_HIP_TRY(tb_launch_dp(a, launches[c].second, dp_st[h]));
        GBSW_HIP_TRY(hipEventRecord(b->tb_events[2 * c], dp_st[h]));
        GBSW_HIP_TRY(hipStreamWaitEvent(walk_st[h], b->tb_events[2 * c], 0));
        GBSW_HIP_TRY(tb_launch_walk(a, launches[c].second, walk_st[h]));
        GBSW_HIP_TRY(hipEventRecord(b->tb_events[2 * c + 1], walk_st[h]));
    }
    for (size_t c = launches.size() >= 2 ? launches.size() - 2 : 0; c < launches.size(); ++c)
        GBSW_HIP_TRY(hipStreamWaitEvent(st, b->tb_events[2 * c + 1], 0));
    GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, st)); // device_ms: the path kernels
    GBSW_HIP_TRY(hipMemcpyAsync(nruns, d_nruns, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    if (path_score)
        GBSW_HIP_TRY(hipMemcpyAsync(path_score, d_score, sizeof(int32_t) * (size_t)n, hipMemcpyDeviceToHost, st));
    GBSW_HIP_TRY(hipStreamSynchronize(st));
    for (int64_t p = 0; p < n; ++p) cigar_offsets[p + 1] = cigar_offsets[p] + nruns[p];
    const int64_t total = cigar_offsets[n];
    if (device_ms) {
        float ms = 0.f;
        GBSW_HIP_TRY(hipEventElapsedTime(&ms, ex.ev_start, ex.ev_stop));
        *device_ms = ms;
    }
    if (total > capacity) {
        return fail(GBSW_E_RANGE, "cigar capacity " + std::to_string(capacity) + " < " + std::to_string(total) +
                                      " words needed (cigar_offsets holds the layout)");
    }
    GBSW_HIP_TRY(b->tb_dst.reserve(sizeof(uint32_t) * (size_t)std::max<int64_t>(1, total)));
    // the offsets through pinned memory (the order and dir_off staging is free now)
    std::memcpy(order, cigar_offsets, sizeof(int64_t) * (size_t)(n + 1));
    GBSW_HIP_TRY(hipMemcpyAsync(d_dst_off, order, sizeof(int64_t) * (size_t)(n + 1), hipMemcpyHostToDevice, st));
    GBSW_HIP_TRY(tb_compact(a.runs, d_cap, d_nruns, d_dst_off, b->tb_dst.as<uint32_t>(), n, st));
    if (total)
        GBSW_HIP_TRY(hipMemcpyAsync(cigar, b->tb_dst.ptr, sizeof(uint32_t) * (size_t)total, hipMemcpyDeviceToHost, st));
    GBSW_HIP_TRY(hipStreamSynchronize(st));
    return GBSW_OK;
}

int gbsw_batch_results(gbsw_batch* b, int32_t* score, int32_t* ref_begin, int32_t* ref_end,
                       int32_t* query_begin, int32_t* query_end)
{
    if (!b) return fail(GBSW_E_ARGUMENT, "batch is null");
    GBSW_HIP_TRY(hipSetDevice(b->ctx->device));
    int32_t* dst[5] = {score, ref_begin, ref_end, query_begin, query_end};
    for (int f = 0; f < 5; ++f)
        if (dst[f] && b->job.n)
            GBSW_HIP_TRY(hipMemcpy(dst[f], b->out.as<int32_t>() + (size_t)f * (size_t)b->job.n,
                                   sizeof(int32_t) * (size_t)b->job.n, hipMemcpyDeviceToHost));
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
    const LaunchRec& L = b->launches[(size_t)index];
    out->kind = L.kind;
    out->columns_per_lane = L.K;
    out->workgroups = L.grid;
    out->pairs = L.pairs;
    out->cells = L.cells;
    out->ms = L.ms;
    return GBSW_OK;
}

int gbsw_batch_autotune(gbsw_batch* b, const gbsw_scoring* sc, int32_t repeats, int32_t* model, double* ms)
{
    if (!b || !sc) return fail(GBSW_E_ARGUMENT, "null argument");
    if (repeats < 1) repeats = 1;
    gbsw_context* ctx = b->ctx;
    GBSW_HIP_TRY(hipSetDevice(ctx->device));
    Job& job = b->job;
    auto use_model = [&](int m) -> int {
        GBSW_HIP_TRY(hipMemcpy(b->bucket_of_len.ptr, bucket_table(job.alpha, m).data(), 65536, hipMemcpyHostToDevice));
        job.kmodel = m;
        count_buckets(job, b->len_ref.data(), b->len_qry.data());
        b->graph.reset(); // its launches hold the previous map's bucket slices
        return GBSW_OK;
    };
    // the batch's current model first; another one must beat it by more than
    // kTuneMargin (run-to-run noise of a batch is about 1 %)
    constexpr double kTuneMargin = 0.015;
    const int current = job.kmodel;
    int best = -1;
    double best_ms = 1e300;
    std::vector<const std::vector<uint8_t>*> tried;
    for (int i = 0; i < kGeometryModels; ++i) {
        const int m = i == 0 ? current : (i <= current ? i - 1 : i);
        // models whose length -> geometry map is one already timed add nothing
        const std::vector<uint8_t>* t = &bucket_table(job.alpha, m);
        bool seen = false;
        for (const auto* u : tried) seen = seen || *u == *t;
        if (seen) continue;
        tried.push_back(t);
        int rc = use_model(m);
        double m_ms = 1e300;
        for (int r = 0; !rc && r < repeats + 1; ++r) {
            rc = gbsw_batch_align(b, sc, GBSW_BATCHED);
            if (!rc && r > 0) m_ms = std::min(m_ms, b->stats.device_ms); // the first run captures the graph
        }
        if (rc) return rc;
        if (best < 0 || m_ms < best_ms * (1.0 - kTuneMargin)) { best_ms = m_ms; best = m; }
    }
    int rc = use_model(best);
    if (rc) return rc;
    ctx->ex.kmodel = best; // batches created afterwards start from it
    if (model) *model = best;
    if (ms) *ms = best_ms;
    return GBSW_OK;
}

int gbsw_align(gbsw_context* ctx, const gbsw_scoring* sc, int32_t mode, int64_t n, const char* ref,
               const uint64_t* ref_off, const char* query, const uint64_t* query_off, int32_t* score,
               int32_t* ref_begin, int32_t* ref_end, int32_t* query_begin, int32_t* query_end)
{
    if (!sc || !ctx) return fail(GBSW_E_ARGUMENT, "null argument");
    // the context's one-shot batch: its device buffers are kept and only grow, so
    // repeated calls allocate and free nothing (a free would also wait for the
    // whole device)
    if (!ctx->oneshot) {
        ctx->oneshot = new gbsw_batch();
        ctx->oneshot->one_shot = true;
    }
    gbsw_batch* b = ctx->oneshot;
    int rc = batch_fill(b, ctx, ctx->ex, sc->alphabet, n, ref, ref_off, query, query_off);
    if (!rc) rc = gbsw_batch_align(b, sc, mode);
    if (!rc) rc = gbsw_batch_results(b, score, ref_begin, ref_end, query_begin, query_end);
    return rc;
}

int gbsw_align_noalloc(gbsw_context* ctx, const gbsw_scoring* sc, int32_t mode, int64_t n, const char* ref,
                       const uint64_t* ref_off, const char* query, const uint64_t* query_off, int32_t* score,
                       int32_t* ref_begin, int32_t* ref_end, int32_t* query_begin, int32_t* query_end)
{
    if (!sc || !ctx) return fail(GBSW_E_ARGUMENT, "null argument");
    if (n < 0 || (n > 0 && (!ref_off || !query_off))) return fail(GBSW_E_ARGUMENT, "null offsets or negative pair count");
    GBSW_HIP_TRY(hipSetDevice(ctx->device));
    const uint64_t rb = n ? ref_off[n] - ref_off[0] : 0;
    const uint64_t qb = n ? query_off[n] - query_off[0] : 0;
    int rc = noalloc_carve(ctx, n, rb, qb);
    gbsw_batch* b = ctx->nbatch;
    if (!rc) rc = batch_fill(b, ctx, ctx->nx, sc->alphabet, n, ref, ref_off, query, query_off);
    if (!rc) rc = batch_align_on(b, ctx->nx, sc, mode);
    if (!rc) rc = gbsw_batch_results(b, score, ref_begin, ref_end, query_begin, query_end);
    return rc;
}

} // extern "C"
