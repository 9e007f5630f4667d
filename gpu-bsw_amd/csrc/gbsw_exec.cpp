// gbsw_exec.cpp -- the alignment pipeline of one job: geometry choice, length
// buckets, device binning, kernel launches on concurrent streams, the int32
// recompute of saturated pairs, and timing.
//
// Replaces the per-chunk body of driver.cpp:88-221 (pack, two-stream copies,
// forward launch, host newMin, reverse launch, copies back).
#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>

#include "gbsw_internal.hpp"

namespace gbsw {

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }
int fail(int code, const std::string& msg)
{
    g_err = msg;
    return code;
}
const char* last_error_text() { return g_err.c_str(); }

void DeviceBuffer::release()
{
    if (ptr && owned) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    owned = true;
}

hipError_t DeviceBuffer::reserve(size_t b)
{
    if (b <= bytes) return hipSuccess;
    if (!owned) return hipErrorOutOfMemory; // a view never grows
    release();
    hipError_t e = hipMalloc(&ptr, b);
    if (e == hipSuccess) bytes = b;
    else ptr = nullptr;
    return e;
}

void PinnedBuffer::release()
{
    if (ptr) (void)hipHostFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

hipError_t PinnedBuffer::reserve(size_t b)
{
    if (b <= bytes) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc(&ptr, b, hipHostMallocDefault);
    if (e == hipSuccess) bytes = b;
    else ptr = nullptr;
    return e;
}

namespace {

// Defaults of the per-context options.  GBSW_F16=0 keeps every bucket on the
// int16 kernels; GBSW_STREAMS=1 serialises the buckets (per-kernel timing and
// counters).
bool default_f16()
{
    const char* e = std::getenv("GBSW_F16");
    return !(e && std::atoi(e) == 0);
}

bool default_graphs()
{
    const char* e = std::getenv("GBSW_GRAPHS");
    return !(e && std::atoi(e) == 0);
}

bool default_split()
{
    const char* e = std::getenv("GBSW_SPLIT");
    return !(e && std::atoi(e) == 0);
}

int default_kmodel()
{
    const char* e = std::getenv("GBSW_KMODEL");
    const int m = e ? std::atoi(e) : kModelOccupancy;
    return (m >= 0 && m < kGeometryModels) ? m : kModelOccupancy;
}

int default_streams()
{
    const char* e = std::getenv("GBSW_STREAMS");
    const int v = e ? std::atoi(e) : kDefaultStreams;
    return std::max(1, std::min(kAuxStreams + 1, v));
}

// LDS bytes of a workgroup; bnd_bytes = 4 (packed int16) or 8 (int32 kernel).
int lds_for(int K, int alpha, bool p8, int bnd_bytes = 4)
{
    return lds_layout(prof_row_bytes(K, p8), alpha, bnd_bytes).total;
}

int grid_for(const Exec& ex, int kind, int K, int lds, int64_t nquads)
{
    int occ = kernel_occupancy(kind, K, lds);
    if (occ < 1) occ = 1;
    const int64_t full = (int64_t)occ * ex.num_cu;
    return (int)std::max<int64_t>(1, std::min<int64_t>(full, nquads));
}

// Launch one kernel between two timing events.
int launch_timed(Exec& ex, hipStream_t stream, int kind, int K, int grid, int lds, KernelArgs& a, void* bnd)
{
    const size_t slot = ex.launches.size();
    a.bnd = bnd;
    while (ex.kev.size() < 2 * (slot + 1)) {
        hipEvent_t e;
        GBSW_HIP_TRY(hipEventCreate(&e));
        ex.kev.push_back(e);
    }
    GBSW_HIP_TRY(hipEventRecord(ex.kev[2 * slot], stream));
    GBSW_HIP_TRY(kernel_launch(kind, K, grid, lds, stream, a));
    GBSW_HIP_TRY(hipEventRecord(ex.kev[2 * slot + 1], stream));
    ex.launches.push_back({kind, K, grid, 0, 0, 0.0});
    return GBSW_OK;
}

KernelArgs base_args(const Exec& ex, const Job& job, const gbsw_scoring* sc)
{
    KernelArgs a;
    std::memset(&a, 0, sizeof(a));
    a.ref = job.ref;
    a.ref_off = job.ref_off;
    a.qry = job.qry;
    a.qry_off = job.qry_off;
    a.n = job.n;
    a.out = job.out;
    a.status = ex.status.as<int32_t>();
    a.redo = ex.redo.as<int32_t>();
    a.lut = ex.lut.as<uint8_t>();
    a.sub = ex.sub.as<int16_t>();
    a.colmax = ex.colmax.as<int16_t>();
    a.alpha = job.alpha;
    a.gap_min = std::min(-(int32_t)sc->gap_open, -(int32_t)sc->gap_extend);
    a.gap_open = sc->gap_open;
    a.gap_extend = sc->gap_extend;
    a.sat_limit = 32767 - ex.max_sub;
    a.asym = table_is_asymmetric(sc, job) ? 1 : 0;
    a.bnd_rows = std::max(1, job.max_rows);
    return a;
}

} // namespace

int Exec::init(int dev, int ncu)
{
    device = dev;
    num_cu = ncu;
    streams = default_streams();
    f16 = default_f16();
    split = default_split();
    graphs = default_graphs();
    kmodel = default_kmodel();
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ev_start);
    if (e == hipSuccess) e = hipEventCreate(&ev_stop);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming);
    if (e == hipSuccess && ensure_streams(streams) != GBSW_OK) e = hipErrorOutOfMemory;
    if (e == hipSuccess) e = status.reserve(sizeof(int32_t) * kStatusWords);
    if (e == hipSuccess) e = status_host.reserve(sizeof(int32_t) * kStatusWords);
    if (e == hipSuccess) e = tables_host.reserve(4096);
    if (e == hipSuccess) e = queue.reserve(sizeof(uint32_t) * 64);
    if (e == hipSuccess) e = sub.reserve(sizeof(int16_t) * 1024);
    if (e == hipSuccess) e = lut.reserve(256);
    if (e == hipSuccess) e = colmax.reserve(sizeof(int16_t) * 32);
    if (e != hipSuccess) return fail(GBSW_E_DEVICE, std::string("stream/scratch setup: ") + hipGetErrorString(e));
    return GBSW_OK;
}

int Exec::init_views(int dev, int ncu)
{
    device = dev;
    num_cu = ncu;
    streams = default_streams();
    f16 = default_f16();
    split = default_split();
    graphs = false;
    fixed = true;
    kmodel = default_kmodel();
    GBSW_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    GBSW_HIP_TRY(hipEventCreate(&ev_start));
    GBSW_HIP_TRY(hipEventCreate(&ev_stop));
    GBSW_HIP_TRY(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
    // every stream a job may use, and the launch events, now: a job then creates nothing
    if (int rc = ensure_streams(kAuxStreams + 1)) return rc;
    while (kev.size() < 2 * (size_t)kNoallocLaunchEvents) {
        hipEvent_t e;
        GBSW_HIP_TRY(hipEventCreate(&e));
        kev.push_back(e);
    }
    GBSW_HIP_TRY(status_host.reserve(sizeof(int32_t) * kStatusWords));
    GBSW_HIP_TRY(tables_host.reserve(4096));
    return GBSW_OK;
}

int Exec::ensure_streams(int n)
{
    for (int i = 0; i < std::min(kAuxStreams, n - 1); ++i) {
        if (aux[i]) continue;
        GBSW_HIP_TRY(hipStreamCreateWithFlags(&aux[i], hipStreamNonBlocking));
        GBSW_HIP_TRY(hipEventCreateWithFlags(&ev_join[i], hipEventDisableTiming));
    }
    return GBSW_OK;
}

void Exec::destroy()
{
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < kAuxStreams; ++i)
        if (aux[i]) (void)hipStreamSynchronize(aux[i]);
    for (hipEvent_t e : kev) (void)hipEventDestroy(e);
    kev.clear();
    if (ev_start) (void)hipEventDestroy(ev_start);
    if (ev_stop) (void)hipEventDestroy(ev_stop);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    for (int i = 0; i < kAuxStreams; ++i) {
        if (ev_join[i]) (void)hipEventDestroy(ev_join[i]);
        if (aux[i]) (void)hipStreamDestroy(aux[i]);
        ev_join[i] = nullptr;
        aux[i] = nullptr;
    }
    bnd.release(); queue.release(); status.release(); sub.release(); lut.release(); colmax.release(); redo.release();
    split_buf.release();
    status_host.release();
    tables_host.release();
    if (stream) (void)hipStreamDestroy(stream);
    stream = nullptr;
    ev_start = ev_stop = ev_fork = nullptr;
}

int max_k_for(int alpha, bool p8)
{
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    int best = ks[0];
    for (int i = 0; i < nk; ++i)
        if (lds_for(ks[i], alpha, p8) <= 65536) best = ks[i];
    return best;
}

double occupancy_threshold(int model)
{
    static const double thr[kGeometryModels] = {0.0, 2.5, 2.0, 3.0, 4.0, 1.5};
    return (model > 0 && model < kGeometryModels) ? thr[model] : 0.0;
}

// Share of the VALU issue rate a geometry reaches at its occupancy: the large
// DNA buckets issue 0.96 VALU instructions per CU per cycle at 4 waves per SIMD,
// 0.82 at 2.75, 0.78 at 2 and 0.62 at 1.5 (rocprofv3 counters), so
// min(1, waves per SIMD / threshold), threshold 2.5 by default.
double occupancy_factor(int K, int alpha, double threshold)
{
    static std::mutex mu;
    static std::map<int, int> cache; // workgroups per CU
    int wg_per_cu = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        const int key = K * 64 + alpha;
        auto it = cache.find(key);
        if (it != cache.end()) {
            wg_per_cu = it->second;
        } else {
            // one wave each; a small alphabet's usual tables (exact f16 high
            // bytes) run the f16 kernels on a byte profile
            const bool p8 = byte_profile_alphabet(alpha);
            wg_per_cu = kernel_occupancy(with_profile(kBatched, p8), K, lds_for(K, alpha, p8));
            if (!p8 && kernel_has(kBatchedF16, K))
                wg_per_cu = std::max(wg_per_cu, kernel_occupancy(with_profile(kBatchedF16, true), K, lds_for(K, alpha, true)));
            cache.emplace(key, wg_per_cu);
        }
    }
    return (wg_per_cu > 0 && threshold > 0) ? std::min(1.0, wg_per_cu / 4.0 / threshold) : 1.0;
}

int k_for_cols(int cols, int alpha, int model)
{
    const int kmax = max_k_for(alpha, byte_profile_alphabet(alpha));
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    int best = ks[0];
    double best_cost = 1e300;
    for (int i = 0; i < nk; ++i) {
        const int K = ks[i];
        if (K > kmax) break;
        const int64_t strips = std::max<int64_t>(1, (cols + kVirtPerPair * K - 1) / (kVirtPerPair * K));
        // padded columns and the per-step cost, per strip
        double cost = (double)(strips * (K + 2) * 4 + (strips - 1));
        if (model != kModelClassic) cost /= occupancy_factor(K, alpha, occupancy_threshold(model));
        if (cost < best_cost) { best_cost = cost; best = K; }
    }
    return best;
}

int bucket_ids()
{
    int nk = 0;
    kernel_k_list(&nk);
    return 3 * nk;
}

StripWidths strip_widths()
{
    StripWidths sw;
    std::memset(&sw, 0, sizeof(sw));
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    for (int b = 0; b < std::min(64, bucket_ids()); ++b) sw.w[b] = (uint16_t)(kVirtPerPair * ks[b % nk]);
    return sw;
}

const std::vector<uint8_t>& bucket_table(int alpha, int model)
{
    static std::mutex mu;
    static std::map<int, std::vector<uint8_t>> cache;
    std::lock_guard<std::mutex> lk(mu);
    const int key = alpha * 64 + model;
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    // strip counts, hence the choice, only change at multiples of 32 columns:
    // the lengths (32(j-1), 32j] share the choice made for 32j
    std::vector<int> kblk(65536 / 32 + 1);
    for (size_t j = 0; j < kblk.size(); ++j) kblk[j] = k_for_cols((int)(32 * j), alpha, model);
    std::vector<uint8_t> t(65536);
    for (int m = 0; m < 65536; ++m) {
        const int K = kblk[(size_t)((m + 31) / 32)];
        int idx = 0;
        while (idx < nk && ks[idx] != K) ++idx;
        // pairs spanning several strips get buckets of their own, so a single-strip
        // bucket bounds its scores by 32 x K columns; those up to kMultiShortCols
        // columns apart from longer ones, whose bound (with DNA scores up to 3)
        // still allows the f16 kernels
        const bool multi = m > kVirtPerPair * K;
        const int cls = !multi ? 0 : m <= kMultiShortCols ? 1 : 2;
        t[(size_t)m] = (uint8_t)(idx + cls * nk);
    }
    return cache.emplace(key, std::move(t)).first->second;
}

void count_buckets(Job& job, const int32_t* len_ref, const int32_t* len_qry)
{
    const std::vector<uint8_t>& tab = bucket_table(job.alpha, job.kmodel);
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    const int nb = bucket_ids();
    std::vector<int64_t> count((size_t)nb, 0), cells((size_t)nb, 0);
    std::vector<int32_t> mcols((size_t)nb, 0);
    for (int64_t p = 0; p < job.n; ++p) {
        const int32_t la = len_ref[p], lb = len_qry[p];
        const int32_t m = std::min(la, lb);
        const int bi = tab[(size_t)std::min<int32_t>(m, 65535)];
        count[(size_t)bi] += 1;
        cells[(size_t)bi] += (int64_t)la * lb;
        mcols[(size_t)bi] = std::max(mcols[(size_t)bi], m);
    }
    job.buckets.clear();
    int64_t at = 0;
    for (int i = 0; i < nb; ++i) {
        if (count[(size_t)i] == 0) continue;
        Bucket bk;
        bk.K = ks[i % nk];
        bk.begin = at;
        bk.count = count[(size_t)i];
        bk.cells = cells[(size_t)i];
        bk.max_cols = mcols[(size_t)i];
        at += bk.count;
        job.buckets.push_back(bk);
    }
}

int batched_grid_cap(const Exec& ex, int K, int alpha)
{
    // the largest occupancy of the geometry's kernels and profile layouts, so
    // that a launch's grid (hence its boundary rows) never exceeds what
    // admission reserved for it
    const bool p8 = byte_profile_alphabet(alpha);
    int occ = kernel_occupancy(with_profile(kBatched, p8), K, lds_for(K, alpha, p8));
    if (kernel_has(kBatchedF16, K)) {
        occ = std::max(occ, kernel_occupancy(kBatchedF16, K, lds_for(K, alpha, false)));
        occ = std::max(occ, kernel_occupancy(with_profile(kBatchedF16, true), K, lds_for(K, alpha, true)));
    }
    return std::max(1, occ) * ex.num_cu;
}

bool f16_high_byte_exact(int32_t s)
{
    // an integer is an f16 value with a zero low byte when it has at most three
    // significant bits (leading one and two mantissa bits) below 2^11
    const int32_t a = s < 0 ? -s : s;
    if (a >= kF16Exact) return false;
    return a == 0 || (a >> __builtin_ctz((uint32_t)a)) < 8;
}

int batched_kind(const Bucket& bk, int32_t max_sub, bool allow_f16)
{
    // every cell value of a pair is at most its score plus one substitution, and
    // the score at most (shorter length) x (best substitution)
    const int64_t bound = (int64_t)bk.max_cols * max_sub + max_sub;
    if (allow_f16 && kernel_has(kBatchedF16, bk.K) && bound < kF16Exact) return kBatchedF16;
    return kBatched;
}

int sort_scratch_bytes(int64_t n, size_t* bytes)
{
    size_t tb = 0;
    GBSW_HIP_TRY(sort_pairs(nullptr, tb, nullptr, nullptr, nullptr, nullptr, (int)std::max<int64_t>(1, n), kSortBits,
                            nullptr));
    *bytes = std::max<size_t>(tb, 16);
    return GBSW_OK;
}

bool table_is_asymmetric(const gbsw_scoring* sc, const Job& job)
{
    if (sc->alphabet != GBSW_PROTEIN) return false;
    for (int r = 0; r < job.alpha; ++r)
        for (int c = 0; c < r; ++c) {
            const int cr = job.code_of ? job.code_of[r] : r, cc = job.code_of ? job.code_of[c] : c;
            if (sc->matrix[cr * 24 + cc] != sc->matrix[cc * 24 + cr]) return true;
        }
    return false;
}

int prepare_scoring(const Job& job, const gbsw_scoring* sc, std::vector<int16_t>& sub, int* max_sub, bool* p8,
                    bool* f16p8)
{
    if (sc->alphabet != job.alphabet) return fail(GBSW_E_ARGUMENT, "scoring alphabet differs from the batch alphabet");
    if (sc->gap_open > 0 || sc->gap_extend > 0) return fail(GBSW_E_SCORING, "gap_open and gap_extend must be <= 0");
    const int A = job.alpha;
    sub.assign((size_t)std::max(1, A * A), 0);
    int mx = 0;
    if (job.alphabet == GBSW_DNA) {
        for (int r = 0; r < A; ++r)
            for (int c = 0; c < A; ++c) sub[(size_t)(r * A + c)] = r == c ? sc->match : sc->mismatch;
        mx = std::max<int>(sc->match, sc->mismatch);
    } else {
        // the caller's table indexed enc(row)*24 + enc(col) (kernel.cpp:785-788),
        // restricted to the residue codes of the batch
        for (int r = 0; r < A; ++r)
            for (int c = 0; c < A; ++c) {
                const int cr = job.code_of ? job.code_of[r] : r, cc = job.code_of ? job.code_of[c] : c;
                const int16_t v = sc->matrix[cr * 24 + cc];
                sub[(size_t)(r * A + c)] = v;
                mx = std::max<int>(mx, v);
            }
    }
    for (int16_t v : sub)
        if (v < -16000 || v > 16000) return fail(GBSW_E_SCORING, "substitution scores must lie in [-16000, 16000]");
    *max_sub = std::max(0, mx);
    int lo = 0;
    for (int16_t v : sub) lo = std::min<int>(lo, v);
    *p8 = byte_profile_alphabet(A) && lo >= -127 && mx <= 127;
    if (f16p8) {
        // the f16 kernels can read a byte profile (the scores' f16 high bytes)
        *f16p8 = true;
        for (int16_t v : sub) *f16p8 = *f16p8 && f16_high_byte_exact(v);
    }
    // result keys hold scores in 23 bits; int16 saturation is handled per pair
    if ((int64_t)job.max_cols * *max_sub >= (1 << 23)) return fail(GBSW_E_RANGE, "possible score exceeds 2^23");
    return GBSW_OK;
}

namespace {

// One length bucket's launch in the batched form.
struct BucketPlan {
    int kind = 0, K = 0, grid = 0, lds = 0, strips = 1;
    size_t bnd_at = 0, split_at = 0; // 32-bit word offsets into ex.bnd / ex.split_buf
};

// Estimated duration of a bucket (relative): cells x (K + 2) / K (padding and
// per-step overhead), +15 % for int16 cells, x2.5 for the split form (its strips
// wait for each other; it issues VALU at 0.3-0.7 of the other kernels' rate).
double bucket_cost(const BucketPlan& b, const Bucket& bk)
{
    const int k = b.kind & ~kByteProfile;
    const double f16 = (k == kBatchedF16 || k == kSplitF16) ? 1.0 : 1.15;
    // a split bucket's strips wait on each other, and beside the bulk buckets it
    // runs several times longer than alone (the K = 7 split bucket of the 1M-pair
    // DNA mix: 1.0 ms alone, 6.6 ms beside them); weights 2.5 / 4 / 6 / 8 / 12
    // measured 3,658-3,672 / 3,703-3,708 / 3,715-3,729 / 3,717-3,726 / 3,713 GCUPS
    const double split = (k == kSplit || k == kSplitF16) ? 6.0 : 1.0;
    return (double)bk.cells * (b.K + 2) / b.K * f16 * split;
}

bool plan_is_split(const BucketPlan& b)
{
    const int k = b.kind & ~kByteProfile;
    return k == kSplit || k == kSplitF16;
}

// 32-bit words of a split bucket's bookkeeping: keys (2 x count u64), progress
// words (2 slots per workgroup, u64), done counters (2 x quads), release epochs
// (one per slot), the placement table (kXcds) and the strip gates (count), padded
// to a multiple of 64 (256 bytes) so that the next bucket's 64-bit keys and
// progress words stay 8-byte aligned.
size_t split_words_of(const Bucket& bk, const BucketPlan& b)
{
    const size_t nquads = (size_t)((bk.count + kPairsPerWave - 1) / kPairsPerWave);
    const size_t words = 5 * (size_t)bk.count + 2 * nquads + 3 * (size_t)(2 * b.grid) + kXcds;
    return (words + 63) & ~(size_t)63;
}

// Kernels, grids and scratch of every bucket (host only; reserves the buffers).
int plan_batched(Exec& ex, const Job& job, int max_sub, bool p8, bool f16p8, bool asym, int kclamp, int64_t bnd_rows,
                 std::vector<BucketPlan>& plan, std::vector<size_t>& order)
{
    const size_t nb = job.buckets.size();
    plan.assign(nb, BucketPlan{});
    order.resize(nb);
    // the unsplit kernel of a bucket and its LDS layout: the int16 kernels use
    // the job's profile (p8), the f16 ones a byte profile when the table allows
    // it (f16p8, a third of the LDS of DNA's int16 profile: more workgroups per
    // CU) and else the int16 profile, which a byte-profile alphabet does not get
    const bool allow_f16 = ex.f16 && (f16p8 || !p8);
    auto unsplit = [&](BucketPlan& b, const Bucket& bk) {
        const int k = batched_kind(bk, max_sub, allow_f16);
        const bool bp8 = k == kBatchedF16 ? f16p8 : p8;
        b.kind = with_profile(k, bp8);
        b.lds = lds_for(b.K, job.alpha, bp8);
    };
    for (size_t i = 0; i < nb; ++i) {
        order[i] = i;
        const Bucket& bk = job.buckets[i];
        BucketPlan& b = plan[i];
        b.K = std::min(kclamp, bk.K);
        unsplit(b, bk);
        const bool bp8 = (b.kind & kByteProfile) != 0;
        const int64_t nquads = (bk.count + kPairsPerWave - 1) / kPairsPerWave;
        b.strips = std::max(1, (bk.max_cols + kVirtPerPair * b.K - 1) / (kVirtPerPair * b.K));
        const int skind = with_profile((b.kind & ~kByteProfile) == kBatchedF16 ? kSplitF16 : kSplit, bp8);
        // A split item waits about four chunks for its predecessor's first
        // rows: worth it only when the bucket has too few quads to keep every
        // workgroup busy for two rounds (on the 200k-pair DNA mix: 1765 vs 827
        // GCUPS at 1.4 quads per workgroup; 988 vs 1275 on protein at 2.7).
        const int64_t full = (int64_t)std::max(1, kernel_occupancy(b.kind, b.K, b.lds)) * ex.num_cu;
        // (an asymmetric table may need a third pass per pair, which the split
        // form's strip items do not schedule)
        if (ex.split && !ex.split_misplaced && !asym && b.strips > 1 && nquads < 2 * full && kernel_has(skind, b.K)) {
            // every XCD's counter needs workgroups: a multiple of kXcds, no more
            // than the items a pipelined quad can keep busy, and two boundary
            // slots per workgroup
            b.kind = skind;
            const int occ = std::max(1, kernel_occupancy(skind, b.K, b.lds));
            const int64_t busy = (nquads * b.strips + kXcds - 1) / kXcds * kXcds;
            b.grid = (int)std::max<int64_t>(kXcds, std::min<int64_t>(occ * ex.num_cu / kXcds * kXcds, busy));
        } else {
            b.grid = grid_for(ex, b.kind, b.K, b.lds, nquads);
        }
    }
    // boundary rows and split bookkeeping (keys: 2 x count u64, prog: slots u64,
    // done: 2 x quads u32, rel: slots u32, xcc, gate: count u32) in 32-bit words
    size_t bnd_words = 0, split_words = 0;
    auto sizes = [&]() {
        bnd_words = split_words = 0;
        for (size_t i = 0; i < nb; ++i) {
            const Bucket& bk = job.buckets[i];
            BucketPlan& b = plan[i];
            b.bnd_at = bnd_words;
            if (plan_is_split(b)) {
                bnd_words += (size_t)2 * b.grid * kPairsPerWave * (size_t)bnd_rows;
                b.split_at = split_words;
                split_words += split_words_of(bk, b);
            } else if (b.strips > 1) {
                // (a bucket whose pairs all fit one strip hands no boundary rows on)
                bnd_words += (size_t)b.grid * kPairsPerWave * (size_t)bnd_rows;
            }
        }
    };
    sizes();
    if (ex.fixed) {
        // preallocated buffers (a serving slot): admission reserved the unsplit
        // boundary rows; split buckets go back to the unsplit form, widest first,
        // until everything fits
        while (bnd_words * sizeof(uint32_t) > ex.bnd.bytes || split_words * sizeof(uint32_t) > ex.split_buf.bytes) {
            size_t w = nb;
            for (size_t i = 0; i < nb; ++i)
                if (plan_is_split(plan[i]) && (w == nb || plan[i].grid > plan[w].grid)) w = i;
            if (w == nb) break;
            unsplit(plan[w], job.buckets[w]);
            plan[w].grid = grid_for(ex, plan[w].kind, plan[w].K, plan[w].lds,
                                    (job.buckets[w].count + kPairsPerWave - 1) / kPairsPerWave);
            sizes();
        }
        // then the widest multi-strip grids shrink (a persistent grid of any size
        // drains its queue) until the boundary rows fit
        while (bnd_words * sizeof(uint32_t) > ex.bnd.bytes) {
            size_t w = nb;
            for (size_t i = 0; i < nb; ++i)
                if (!plan_is_split(plan[i]) && plan[i].strips > 1 && plan[i].grid > 1 &&
                    (w == nb || plan[i].grid > plan[w].grid))
                    w = i;
            if (w == nb) return fail(GBSW_E_RANGE, "strip boundary rows exceed the preallocated scratch");
            plan[w].grid = std::max(1, plan[w].grid / 2);
            sizes();
        }
    }
    // Buckets run concurrently on the streams, the ones whose single work item
    // takes longest (largest K) first, so a bucket with few long pairs overlaps
    // the bulk instead of running alone at the end (ordering by estimated total
    // instead measured 4 % slower on the 1M-pair DNA mix); record_batched puts
    // each on the least-loaded stream.
    std::sort(order.begin(), order.end(), [&](size_t x, size_t y) { return job.buckets[x].K > job.buckets[y].K; });
    if (ex.queue.bytes < sizeof(uint32_t) * kXcds * std::max<size_t>(nb, 1))
        GBSW_HIP_TRY(ex.queue.reserve(sizeof(uint32_t) * kXcds * std::max<size_t>(nb, 1)));
    if (split_words) GBSW_HIP_TRY(ex.split_buf.reserve(split_words * sizeof(uint32_t)));
    GBSW_HIP_TRY(ex.bnd.reserve(std::max<size_t>(bnd_words, 1) * sizeof(uint32_t)));
    return GBSW_OK;
}

// Stream work of the batched form (recorded eagerly or into a graph).
int record_batched(Exec& ex, const Job& job, KernelArgs a, int max_sub, const std::vector<BucketPlan>& plan,
                   const std::vector<size_t>& order, size_t sub_bytes)
{
    hipStream_t st0 = ex.stream;
    const uint8_t* th = ex.tables_host.as<uint8_t>();
    GBSW_HIP_TRY(hipMemcpyAsync(ex.sub.ptr, th, sub_bytes, hipMemcpyHostToDevice, st0));
    GBSW_HIP_TRY(hipMemcpyAsync(ex.lut.ptr, th + 2048, 256, hipMemcpyHostToDevice, st0));
    GBSW_HIP_TRY(hipMemcpyAsync(ex.colmax.ptr, th + 2304, sizeof(int16_t) * (size_t)job.alpha, hipMemcpyHostToDevice,
                                st0));
    GBSW_HIP_TRY(zero_u32(ex.status.as<uint32_t>(), kStatusWords, st0));
    // device binning: keys, then a stable radix sort into the work order
    GBSW_HIP_TRY(bin_keys(job.ref_off, job.qry_off, job.n, job.bucket_of_len, strip_widths(), job.keys, job.vals, st0));
    size_t tb = job.sort_temp_bytes;
    GBSW_HIP_TRY(sort_pairs(job.sort_temp, tb, job.keys, job.keys_sorted, job.vals, job.order, (int)job.n, kSortBits,
                            st0));
    const size_t nb = plan.size();
    GBSW_HIP_TRY(zero_u32(ex.queue.as<uint32_t>(), (int)(kXcds * nb), st0));
    size_t split_words = 0;
    for (size_t i = 0; i < nb; ++i)
        if (plan_is_split(plan[i]))
            split_words = std::max(split_words, plan[i].split_at + split_words_of(job.buckets[i], plan[i]));
    if (split_words) GBSW_HIP_TRY(zero_u32(ex.split_buf.as<uint32_t>(), (int)split_words, st0));
    const int ns = std::max(1, std::min(kAuxStreams + 1, ex.streams));
    GBSW_HIP_TRY(hipEventRecord(ex.ev_fork, st0));
    for (int i = 0; i < ns - 1; ++i) GBSW_HIP_TRY(hipStreamWaitEvent(ex.aux[i], ex.ev_fork, 0));
    // each bucket (in the plan's order) goes to the stream with the least
    // estimated work queued so far (bucket_cost): a short bucket then does not
    // wait behind a long one and run alone at the end
    std::vector<double> load((size_t)ns, 0.0);
    for (size_t j = 0; j < order.size(); ++j) {
        const size_t i = order[j];
        size_t si = 0;
        for (size_t q = 1; q < load.size(); ++q)
            if (load[q] < load[si]) si = q;
        load[si] += bucket_cost(plan[i], job.buckets[i]);
        const Bucket& bk = job.buckets[i];
        const BucketPlan& b = plan[i];
        a.order = job.order + bk.begin;
        a.order_n = bk.count;
        a.queue = ex.queue.as<uint32_t>() + kXcds * i;
        const int base_kind = b.kind & ~kByteProfile;
        a.sat_limit = (base_kind == kBatchedF16 || base_kind == kSplitF16) ? kF16Exact - max_sub : 32767 - max_sub;
        std::memset(&a.split, 0, sizeof(a.split));
        if (plan_is_split(b)) {
            uint32_t* w = ex.split_buf.as<uint32_t>() + b.split_at;
            const int64_t nquads = (bk.count + kPairsPerWave - 1) / kPairsPerWave;
            const int64_t nslots = 2 * (int64_t)b.grid;
            a.split.keys = (uint64_t*)w;
            a.split.prog = (uint64_t*)(w + 4 * bk.count);
            a.split.done = w + 4 * bk.count + 2 * nslots;
            a.split.rel = a.split.done + 2 * nquads;
            a.split.xcc = a.split.rel + nslots;
            a.split.gate = a.split.xcc + kXcds;
            if (((uintptr_t)a.split.keys | (uintptr_t)a.split.prog) & 7)
                return fail(GBSW_E_INTERNAL, "split bookkeeping of a bucket is not 8-byte aligned");
            a.split.strips = b.strips;
            a.split.slots = (int32_t)(nslots / kXcds);
        }
        hipStream_t st = si == 0 ? st0 : ex.aux[si - 1];
        int rc = launch_timed(ex, st, b.kind, b.K, b.grid, b.lds, a, ex.bnd.as<uint32_t>() + b.bnd_at);
        if (rc) return rc;
        ex.launches.back().pairs = bk.count;
        ex.launches.back().cells = bk.cells;
    }
    for (int i = 0; i < ns - 1; ++i) {
        GBSW_HIP_TRY(hipEventRecord(ex.ev_join[i], ex.aux[i]));
        GBSW_HIP_TRY(hipStreamWaitEvent(st0, ex.ev_join[i], 0));
    }
    GBSW_HIP_TRY(hipMemcpyAsync(ex.status_host.ptr, ex.status.ptr, sizeof(int32_t) * kStatusWords,
                                hipMemcpyDeviceToHost, st0));
    return GBSW_OK;
}

// What a captured graph depends on besides the batch's bytes: the scoring, the
// options, the addresses of the context buffers it writes, and each bucket's
// slice of the sort order (baked into the captured launch arguments; another
// length -> geometry map of the same batch moves the slices while the plans may
// stay alike).
std::string graph_key(const Exec& ex, const Job& job, const gbsw_scoring* sc, const std::vector<BucketPlan>& plan)
{
    // DNA reads the first four scores only
    const size_t sc_bytes = sc->alphabet == GBSW_DNA ? offsetof(gbsw_scoring, matrix) : sizeof(*sc);
    std::string k((const char*)sc, sc_bytes);
    const void* ptrs[] = {ex.bnd.ptr,   ex.queue.ptr,          ex.split_buf.ptr, ex.redo.ptr,       ex.status.ptr,
                          ex.sub.ptr,   ex.lut.ptr,            ex.colmax.ptr,    ex.tables_host.ptr, ex.status_host.ptr,
                          ex.stream};
    k.append((const char*)ptrs, sizeof(ptrs));
    const int opts[] = {ex.streams, ex.f16 ? 1 : 0, (ex.split && !ex.split_misplaced) ? 1 : 0, (int)plan.size()};
    k.append((const char*)opts, sizeof(opts));
    const int64_t jf[] = {job.kmodel, job.n, (int64_t)(uintptr_t)job.order};
    k.append((const char*)jf, sizeof(jf));
    for (size_t i = 0; i < plan.size(); ++i) {
        const BucketPlan& b = plan[i];
        const Bucket& bk = job.buckets[i];
        const int64_t f[] = {b.kind, b.K, b.grid, b.lds, b.strips, (int64_t)b.bnd_at, (int64_t)b.split_at,
                             bk.begin, bk.count, bk.max_cols};
        k.append((const char*)f, sizeof(f));
    }
    return k;
}

} // namespace

int enqueue_job(Exec& ex, const Job& job, const gbsw_scoring* sc, int mode)
{
    ex.launches.clear();
    ex.graph_state = 0;
    std::vector<int16_t> sub;
    int max_sub = 0;
    bool p8 = false, f16p8 = false;
    int rc = prepare_scoring(job, sc, sub, &max_sub, &p8, &f16p8);
    if (rc) return rc;
    ex.max_sub = max_sub;
    ex.p8 = p8;
    // geometries were chosen for the byte profile of a large alphabet; scores
    // outside int8 take the int16 profile, in the widest geometry that fits
    const int kclamp = (byte_profile_alphabet(job.alpha) && !p8) ? max_k_for(job.alpha, false) : 1 << 30;
    hipStream_t st0 = ex.stream;
    if (job.n == 0) {
        GBSW_HIP_TRY(hipEventRecord(ex.ev_start, st0));
        GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, st0));
        std::memset(ex.status_host.ptr, 0, sizeof(int32_t) * kStatusWords);
        return GBSW_OK;
    }
    std::vector<int16_t> colmax((size_t)job.alpha, 0);
    for (int c = 0; c < job.alpha; ++c)
        for (int r = 0; r < job.alpha; ++r)
            colmax[(size_t)c] = std::max<int16_t>(colmax[(size_t)c], sub[(size_t)(r * job.alpha + c)]);
    // small tables through pinned staging (the previous job of this Exec has
    // completed, so its staging is free): the copies stay asynchronous
    uint8_t* th = ex.tables_host.as<uint8_t>();
    std::memcpy(th, sub.data(), sizeof(int16_t) * sub.size());
    std::memcpy(th + 2048, job.lut_host, 256);
    std::memcpy(th + 2304, colmax.data(), sizeof(int16_t) * colmax.size());
    GBSW_HIP_TRY(ex.redo.reserve(sizeof(int32_t) * (size_t)std::max<int64_t>(1, job.n)));
    KernelArgs a = base_args(ex, job, sc);
    if (mode == GBSW_PLAIN) {
        GBSW_HIP_TRY(hipEventRecord(ex.ev_start, st0));
        GBSW_HIP_TRY(hipMemcpyAsync(ex.sub.ptr, th, sizeof(int16_t) * sub.size(), hipMemcpyHostToDevice, st0));
        GBSW_HIP_TRY(hipMemcpyAsync(ex.lut.ptr, th + 2048, 256, hipMemcpyHostToDevice, st0));
        GBSW_HIP_TRY(hipMemcpyAsync(ex.colmax.ptr, th + 2304, sizeof(int16_t) * colmax.size(), hipMemcpyHostToDevice, st0));
        GBSW_HIP_TRY(zero_u32(ex.status.as<uint32_t>(), kStatusWords, st0));
        // one geometry for the whole batch, sized by its longest shorter sequence;
        // with profiles (DNA) the occupancy model's narrower strips cost the long
        // pairs more than the extra waves gain (793 vs 593 GCUPS on the 1M-pair
        // DNA mix), so it applies to profile-free alphabets only
        const int K = std::min(kclamp, k_for_cols(job.max_cols, job.alpha,
                                                  byte_profile_alphabet(job.alpha) ? job.kmodel : kModelClassic));
        const int lds = lds_for(K, job.alpha, p8);
        const int64_t nquads = (job.n + kPairsPerWave - 1) / kPairsPerWave;
        a.order = nullptr;
        a.order_n = job.n;
        const int kf = with_profile(kPlainForward, p8), kr = with_profile(kPlainReverse, p8);
        int gf = grid_for(ex, kf, K, lds, nquads);
        int gr = grid_for(ex, kr, K, lds, nquads);
        const size_t per_wg = kPairsPerWave * (size_t)a.bnd_rows * sizeof(uint32_t);
        if (ex.fixed && (size_t)std::max(gf, gr) * per_wg > ex.bnd.bytes) {
            // preallocated scratch: fewer workgroups stride over the quads
            const int cap = (int)std::min<size_t>(ex.bnd.bytes / per_wg, INT_MAX);
            if (cap < 1) return fail(GBSW_E_RANGE, "strip boundary rows exceed the preallocated scratch");
            gf = std::min(gf, cap);
            gr = std::min(gr, cap);
        }
        GBSW_HIP_TRY(ex.bnd.reserve((size_t)std::max(gf, gr) * per_wg));
        rc = launch_timed(ex, st0, kf, K, gf, lds, a, ex.bnd.ptr);
        if (rc) return rc;
        ex.launches.back().pairs = job.n;
        ex.launches.back().cells = job.cells;
        rc = launch_timed(ex, st0, kr, K, gr, lds, a, ex.bnd.ptr);
        if (rc) return rc;
        ex.launches.back().pairs = job.n;
        GBSW_HIP_TRY(hipMemcpyAsync(ex.status_host.ptr, ex.status.ptr, sizeof(int32_t) * kStatusWords,
                                    hipMemcpyDeviceToHost, st0));
        GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, st0));
        return GBSW_OK;
    }
    std::vector<BucketPlan> plan;
    std::vector<size_t> order;
    rc = plan_batched(ex, job, max_sub, p8, f16p8, a.asym != 0, kclamp, a.bnd_rows, plan, order);
    if (rc) return rc;
    const size_t sub_bytes = sizeof(int16_t) * sub.size();
    if (!ex.graphs || ex.fixed || !job.graph) {
        GBSW_HIP_TRY(hipEventRecord(ex.ev_start, st0));
        rc = record_batched(ex, job, a, max_sub, plan, order, sub_bytes);
        if (rc) return rc;
        GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, st0));
        return GBSW_OK;
    }
    // hipGraph: the stream work of this batch, scoring and options is captured
    // once and replayed (one launch instead of ~20 API calls); the per-launch
    // records of the capture stand for every replay
    JobGraph& G = *job.graph;
    const std::string key = graph_key(ex, job, sc, plan);
    if (!G.exec || G.key != key) {
        G.reset();
        GBSW_HIP_TRY(hipStreamBeginCapture(st0, hipStreamCaptureModeRelaxed));
        rc = record_batched(ex, job, a, max_sub, plan, order, sub_bytes);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(st0, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (ce != hipSuccess) return fail(GBSW_E_DEVICE, std::string("graph capture: ") + hipGetErrorString(ce));
        const hipError_t ie = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) {
            G.exec = nullptr;
            return fail(GBSW_E_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(ie));
        }
        G.key = key;
        G.launches = ex.launches;
        G.captures += 1;
        ex.graph_state = 1;
    } else {
        ex.graph_state = 2;
    }
    ex.launches = G.launches;
    GBSW_HIP_TRY(hipEventRecord(ex.ev_start, st0));
    GBSW_HIP_TRY(hipGraphLaunch(G.exec, st0));
    GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, st0));
    G.replays += 1;
    return GBSW_OK;
}

int complete_job(Exec& ex, const Job& job, const gbsw_scoring* sc, double* device_ms)
{
    GBSW_HIP_TRY(hipEventSynchronize(ex.ev_stop));
    ex.late_results = false;
    int32_t* status = ex.status_host.as<int32_t>();
    if (split_rerun_needed(status) && !ex.split_misplaced) {
        // a split launch found workgroups of one XCD counter on different XCCs,
        // so its boundary rows may have been read stale: the whole job again,
        // with the split form off from now on (it is only used in the batched form)
        ex.split_misplaced = true;
        set_error("split form disabled: workgroups of one XCD counter ran on different XCCs");
        int rc = enqueue_job(ex, job, sc, GBSW_BATCHED);
        if (!rc) rc = complete_job(ex, job, sc, device_ms);
        ex.late_results = true;
        return rc;
    }
    if (status[kStatusSaturated] > 0) {
        // pairs whose int16 score reached the saturation limit: recompute them
        // (forward + reverse) with int32 cells
        int nk = 0;
        const int* wk = wide_k_list(&nk);
        int K = wk[nk - 1];
        for (int i = 0; i < nk; ++i)
            if (kVirtPerPair * wk[i] >= job.max_cols) { K = wk[i]; break; }
        const int lds = lds_for(K, job.alpha, ex.p8, 8);
        const int64_t count = status[kStatusSaturated];
        KernelArgs a = base_args(ex, job, sc);
        a.order = ex.redo.as<int32_t>();
        a.order_n = count;
        a.sat_limit = INT_MAX;
        a.queue = ex.queue.as<uint32_t>();
        GBSW_HIP_TRY(zero_u32(a.queue, 1, ex.stream));
        const int kw = with_profile(kBatchedWide, ex.p8);
        int grid = grid_for(ex, kw, K, lds, (count + kPairsPerWave - 1) / kPairsPerWave);
        const size_t per_wg = kPairsPerWave * (size_t)a.bnd_rows * 8;
        if (ex.fixed && (size_t)grid * per_wg > ex.bnd.bytes) {
            const int cap = (int)std::min<size_t>(ex.bnd.bytes / per_wg, INT_MAX);
            if (cap < 1) return fail(GBSW_E_RANGE, "int32 boundary rows exceed the preallocated scratch");
            grid = std::min(grid, cap);
        }
        GBSW_HIP_TRY(ex.bnd.reserve((size_t)grid * per_wg));
        int rc = launch_timed(ex, ex.stream, kw, K, grid, lds, a, ex.bnd.ptr);
        if (rc) return rc;
        ex.launches.back().pairs = count;
        GBSW_HIP_TRY(hipMemcpyAsync(ex.status_host.ptr, ex.status.ptr, sizeof(int32_t) * kStatusWords,
                                    hipMemcpyDeviceToHost, ex.stream));
        GBSW_HIP_TRY(hipEventRecord(ex.ev_stop, ex.stream));
        GBSW_HIP_TRY(hipEventSynchronize(ex.ev_stop));
        ex.late_results = true;
    }
    float ms = 0.f;
    GBSW_HIP_TRY(hipEventElapsedTime(&ms, ex.ev_start, ex.ev_stop));
    if (device_ms) *device_ms = ms;
    for (size_t i = 0; i < ex.launches.size(); ++i) {
        // events recorded by a graph replay may not carry times: 0 then
        float t = 0.f;
        if (hipEventElapsedTime(&t, ex.kev[2 * i], ex.kev[2 * i + 1]) != hipSuccess) {
            t = 0.f;
            (void)hipGetLastError();
        }
        ex.launches[i].ms = t;
    }
    if (status[kStatusMisplaced]) return fail(GBSW_E_INTERNAL, "split-form placement check failed after an unsplit rerun");
    // a timed-out wait lets every later wait return at once, after which passes
    // may read boundary rows that are not there: report the timeout, not what
    // follows from it
    if (status[kStatusTimeout]) return fail(GBSW_E_INTERNAL, "a split-form wait timed out");
    if (status[kStatusRevMiss]) return fail(GBSW_E_INTERNAL, "reverse pass did not reach the forward score");
    return GBSW_OK;
}

} // namespace gbsw
