// gbsw_driver.cpp -- gpu_bsw_driver drop-in (include/gpu_bsw_driver.hpp) over the C ABI.
//
// Mirrors the reference driver's contract (driver.cpp:15-243, 245-489;
// utils.cpp:13-28): string vectors in, five pinned `short` arrays out, the
// "Total ..." summary lines on stdout.  The strings are packed into pinned
// staging (as the reference packs into pinned buffers, driver.cpp:113-161, here
// on several threads and kept across calls), then device transfer and both
// passes go through gbsw_align.  Calls of 131,072 pairs or more run as four
// sub-batches alternating between two contexts: while one sub-batch aligns, the
// next is packed and uploaded.
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <algorithm>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gbsw.h"
#include "../../include/gpu_bsw_driver.hpp"

namespace {

thread_local int g_status = 0;
// Large calls run as sub-batches on two contexts (two sets of streams), so the
// upload of one overlaps the alignment of the other while the host packs the next.
constexpr int kLanes = 2;
constexpr size_t kPipelinePairs = 131072; // below this, one sub-batch
constexpr int kSubBatches = 4;
std::mutex g_ctx_mu;
gbsw_context* g_ctx[kLanes] = {};

gbsw_context* shared_context(int lane, int* rc)
{
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (g_ctx[lane]) { *rc = GBSW_OK; return g_ctx[lane]; }
    int dev = 0;
    if (const char* e = std::getenv("GBSW_DEVICE")) dev = std::atoi(e);
    else (void)hipGetDevice(&dev);
    *rc = gbsw_context_create(dev, &g_ctx[lane]);
    return *rc == GBSW_OK ? g_ctx[lane] : nullptr;
}

int32_t driver_mode()
{
    const char* e = std::getenv("GBSW_MODE");
    return (e && std::strcmp(e, "plain") == 0) ? GBSW_PLAIN : GBSW_BATCHED;
}

// initialize_alignments (utils.cpp:13-19): pinned host arrays owned by the caller.
bool allocate_results(gpu_bsw_driver::alignment_results* r, size_t n)
{
    short** f[5] = {&r->ref_begin, &r->ref_end, &r->query_begin, &r->query_end, &r->top_scores};
    const size_t bytes = sizeof(short) * (n ? n : 1);
    for (short** p : f) {
        *p = nullptr;
        if (hipHostMalloc((void**)p, bytes, hipHostMallocDefault) != hipSuccess) return false;
        std::memset(*p, 0, bytes);
    }
    return true;
}

// Pinned host staging, grown on demand and kept for later calls: the upload
// from it runs at the DMA rate, without the runtime's bounce buffer for
// pageable memory.
struct Staging {
    char* p = nullptr;
    size_t cap = 0;
    bool reserve(size_t need)
    {
        if (need <= cap) return true;
        if (p) (void)gbsw_host_free(p);
        p = nullptr;
        cap = 0;
        const size_t grow = std::max(need, (size_t)1 << 20);
        if (gbsw_host_alloc(grow, (void**)&p) != GBSW_OK) return false;
        cap = grow;
        return true;
    }
};
std::mutex g_stage_mu; // the staging sets, held for a whole call
Staging g_stage_ref[kLanes], g_stage_qry[kLanes];

// Offsets of strings v[first .. first + n) relative to the first.
void offsets(const std::vector<std::string>& v, size_t first, std::vector<uint64_t>& off, size_t n)
{
    off.resize(n + 1);
    off[0] = 0;
    for (size_t k = 0; k < n; ++k) off[k + 1] = off[k] + v[first + k].size();
}

// The strings v[first .. first + n) at their offsets in dst, copied by up to 8
// threads (one for small inputs).
void pack(const std::vector<std::string>& v, size_t first, const std::vector<uint64_t>& off, size_t n, char* dst)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const unsigned T = off[n] < ((uint64_t)16 << 20) ? 1u : std::min(8u, hw);
    auto part = [&](unsigned t) {
        for (size_t k = n * t / T; k < n * (t + 1) / T; ++k)
            if (!v[first + k].empty()) std::memcpy(dst + off[k], v[first + k].data(), v[first + k].size());
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < T; ++t) th.emplace_back(part, t);
    part(0);
    for (std::thread& x : th) x.join();
}

[[noreturn]] void die(int rc, const std::string& msg = std::string())
{
    std::fprintf(stderr, "GPUassert: %s (gbsw status %d)\n", msg.empty() ? gbsw_last_error() : msg.c_str(), rc);
    std::exit(1);
}

// Sub-batches of a call: four for large calls (pipelining), and at least as
// many as keep the device footprint of the two sub-batches in flight within
// factor x device memory.  As in the reference (driver.cpp:36, 51-55) a pair is
// priced at the longest ref plus the longest query; here plus kPairBytes of
// offsets, results and binning arrays per pair.
size_t sub_batches(size_t n, size_t max_ref, size_t max_qry, float factor)
{
    constexpr size_t kPairBytes = 96;
    size_t S = n >= kPipelinePairs ? kSubBatches : 1;
    int dev = 0;
    if (const char* e = std::getenv("GBSW_DEVICE")) dev = std::atoi(e);
    else (void)hipGetDevice(&dev);
    uint64_t mem = 0;
    if (n == 0 || !(factor > 0.f) || gbsw_device_memory(dev, &mem) != GBSW_OK) return S;
    const double budget = (double)mem * (double)factor / kLanes;
    const double per_pair = (double)(max_ref + max_qry + kPairBytes);
    const size_t fit = (size_t)std::max(1.0, budget / per_pair);
    return std::max(S, (n + fit - 1) / fit);
}

void run(std::vector<std::string>& reads, std::vector<std::string>& contigs,
         gpu_bsw_driver::alignment_results* alignments, const gbsw_scoring& sc, float factor)
{
    const size_t n = contigs.size(); // driver.cpp:21 pairs reads[k] with contigs[k]
    g_status = GBSW_OK;
    if (!allocate_results(alignments, n)) {
        g_status = GBSW_E_MEMORY;
        if (!std::getenv("GBSW_DRIVER_NOEXIT")) die(g_status);
        return;
    }
    if (reads.size() < n) {
        g_status = GBSW_E_ARGUMENT;
        if (!std::getenv("GBSW_DRIVER_NOEXIT")) {
            std::fprintf(stderr, "GPUassert: fewer reads than contigs\n");
            std::exit(1);
        }
        return;
    }
    size_t max_ref = 0, max_qry = 0;
    for (size_t k = 0; k < n; ++k) {
        max_ref = std::max(max_ref, contigs[k].size());
        max_qry = std::max(max_qry, reads[k].size());
    }
    const auto t0 = std::chrono::high_resolution_clock::now();
    int rc = GBSW_OK;
    std::string err; // a sub-batch's error message (set on its own thread)
    std::vector<int32_t> out(5 * (n ? n : 1));
    {
        std::lock_guard<std::mutex> lk(g_stage_mu);
        // sub-batch k: pairs [n k / S, n (k+1) / S) on lane k % 2; the lane's
        // previous sub-batch has finished before its staging is refilled
        const size_t S = sub_batches(n, max_ref, max_qry, factor);
        std::vector<uint64_t> ro[kLanes], qo[kLanes];
        std::thread worker[kLanes];
        int lane_rc[kLanes] = {GBSW_OK, GBSW_OK};
        std::string lane_err[kLanes];
        const int32_t mode = driver_mode();
        for (size_t k = 0; k < S && rc == GBSW_OK; ++k) {
            const int lane = (int)(k % kLanes);
            if (worker[lane].joinable()) worker[lane].join();
            if (lane_rc[lane] != GBSW_OK) break;
            gbsw_context* ctx = shared_context(lane, &rc);
            if (rc != GBSW_OK) break;
            const size_t p0 = n * k / S, m = n * (k + 1) / S - p0;
            offsets(contigs, p0, ro[lane], m);
            offsets(reads, p0, qo[lane], m);
            Staging& sr = g_stage_ref[lane];
            Staging& sq = g_stage_qry[lane];
            if (!sr.reserve(ro[lane][m] + 1) || !sq.reserve(qo[lane][m] + 1)) {
                rc = GBSW_E_MEMORY;
                break;
            }
            pack(contigs, p0, ro[lane], m, sr.p);
            pack(reads, p0, qo[lane], m, sq.p);
            int32_t* o = out.data() + p0;
            worker[lane] = std::thread([&, lane, ctx, m, o]() {
                lane_rc[lane] = gbsw_align(ctx, &sc, mode, (int64_t)m, g_stage_ref[lane].p, ro[lane].data(),
                                           g_stage_qry[lane].p, qo[lane].data(), o, o + n, o + 2 * n, o + 3 * n,
                                           o + 4 * n);
                if (lane_rc[lane] != GBSW_OK) lane_err[lane] = gbsw_last_error();
            });
        }
        for (std::thread& w : worker)
            if (w.joinable()) w.join();
        for (int lane = 0; lane < kLanes && rc == GBSW_OK; ++lane) {
            rc = lane_rc[lane];
            err = lane_err[lane];
        }
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    g_status = rc;
    if (rc != GBSW_OK) {
        if (!std::getenv("GBSW_DRIVER_NOEXIT")) die(rc, err);
        return;
    }
    // The contract's arrays are `short` (driver.hpp:21-27), whose values the
    // reference would wrap past 32767; here such a call reports GBSW_E_RANGE
    // (exit unless GBSW_DRIVER_NOEXIT) and the arrays hold the values clamped to
    // the short range.  gbsw_align returns the int32 values.
    size_t wide = 0;
    auto to_short = [&](int32_t v) -> short {
        if (v > SHRT_MAX || v < SHRT_MIN) {
            ++wide;
            return (short)(v > SHRT_MAX ? SHRT_MAX : SHRT_MIN);
        }
        return (short)v;
    };
    for (size_t k = 0; k < n; ++k) {
        alignments->top_scores[k] = to_short(out[k]);
        alignments->ref_begin[k] = to_short(out[n + k]);
        alignments->ref_end[k] = to_short(out[2 * n + k]);
        alignments->query_begin[k] = to_short(out[3 * n + k]);
        alignments->query_end[k] = to_short(out[4 * n + k]);
    }
    if (wide) {
        g_status = GBSW_E_RANGE;
        const std::string msg = std::to_string(wide) + " result values exceed the short range of alignment_results";
        if (!std::getenv("GBSW_DRIVER_NOEXIT")) die(g_status, msg);
        std::fprintf(stderr, "gpu_bsw_driver: %s (status %d)\n", msg.c_str(), g_status);
    }
    const std::chrono::duration<double> dt = t1 - t0;
    std::cout << "Total Alignments:" << n << "\n"
              << "Max Reference Size:" << max_ref << "\n"
              << "Max Query Size:" << max_qry << "\n"
              << "Total Execution Time (seconds):" << dt.count() << std::endl;
}

} // namespace

namespace gpu_bsw_driver {

size_t get_tot_gpu_mem(int id)
{
    uint64_t b = 0;
    if (gbsw_device_memory(id, &b) != GBSW_OK) return 0;
    return (size_t)b;
}

void kernel_driver_dna(std::vector<std::string> reads, std::vector<std::string> contigs,
                       alignment_results* alignments, short scores[4], float factor)
{
    gbsw_scoring sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.alphabet = GBSW_DNA;
    sc.match = scores[0];      // driver.cpp:18
    sc.mismatch = scores[1];
    sc.gap_open = scores[2];
    sc.gap_extend = scores[3];
    run(reads, contigs, alignments, sc, factor);
}

void kernel_driver_aa(std::vector<std::string> reads, std::vector<std::string> contigs,
                      alignment_results* alignments, short scoring_matrix[], short openGap, short extendGap,
                      float factor)
{
    gbsw_scoring sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.alphabet = GBSW_PROTEIN;
    sc.gap_open = openGap;
    sc.gap_extend = extendGap;
    std::memcpy(sc.matrix, scoring_matrix, sizeof(sc.matrix)); // SCORE_MAT_SIZE = 576 (kernel.hpp:9)
    run(reads, contigs, alignments, sc, factor);
}

int last_status() { return g_status; }

} // namespace gpu_bsw_driver

// utils.cpp:21-28
void free_alignments(gpu_bsw_driver::alignment_results* alignments)
{
    if (!alignments) return;
    short* f[5] = {alignments->ref_begin, alignments->ref_end, alignments->query_begin, alignments->query_end,
                   alignments->top_scores};
    for (short* p : f)
        if (p) (void)hipHostFree(p);
}
