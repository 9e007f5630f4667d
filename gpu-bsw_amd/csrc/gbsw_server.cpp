// gbsw_server.cpp -- serving loop with continuous batching (gbsw_server_* of
// include/gbsw.h).  The reference has no counterpart: its driver is one blocking
// call per batch (driver.hpp:32-37) that chunks by a memory budget
// (driver.cpp:36, 51-55).  Here the memory budget becomes admission control over
// device-resident micro-batch slots.
//
// Life of a request:
//   submit (caller thread): validate, rebase offsets, summarise lengths per
//     length bucket; copy pageable residues into the pinned arena; upload those
//     of a large request by DMA into the device staging ring; queue it under
//     its scoring class.
//   admit (scheduler thread): when a slot is free, take the scoring class of
//     the oldest pending request and pack its queue, oldest first, into the slot
//     while the slot's pair, residue and strip-boundary capacities hold; a
//     request that does not fit is split at a pair boundary.
//   launch: one host->device copy of the offsets and gather tiles, the gather
//     kernel (staging ring or pinned host memory -> the slot's packed buffers
//     in HBM, alphabet check), after the uploads enqueued so far; the batched
//     pipeline of
//     gbsw_exec.cpp on the slot's streams, one device->host copy of the results.
//   complete: when the slot's event has fired, results are scattered into each
//     request's output arrays; a request whose parts are all done is reported.
// Slots run concurrently, so requests that arrive while others are on the
// device are admitted as soon as a slot frees up and overlap with them.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "gbsw_internal.hpp"

namespace gbsw {
int build_alphabet(int32_t alphabet, const uint8_t* const* bytes, const uint64_t* lens, int nbuf, uint8_t* lut);
}

using namespace gbsw;

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point a, Clock::time_point b)
{
    return std::chrono::duration<double, std::milli>(b - a).count();
}

Clock::time_point trace_epoch()
{
    static const Clock::time_point t0 = Clock::now();
    return t0;
}

constexpr int kMaxBuckets = 64;
// Pinned host memory that kernels read in place: mapped, and coherent so that a
// region rewritten by the host after an earlier read is never served stale.
constexpr unsigned kDeviceReadFlags = hipHostMallocMapped | hipHostMallocCoherent;
constexpr uint64_t kMaxLen = 65000;
// Requests of at least this many residue bytes are uploaded by DMA at submit.
constexpr uint64_t kDmaMinBytes = (uint64_t)256 << 10;

// ---- pinned host blocks handed out by gbsw_host_alloc ----------------------
struct HostBlock {
    size_t bytes;
    const uint8_t* dev;
};
std::mutex g_host_mu;
std::map<uintptr_t, HostBlock> g_host_blocks;

// Device-readable address of [p, p + len) if it lies in one gbsw_host_alloc block.
const uint8_t* registered_device_ptr(const void* p, size_t len)
{
    std::lock_guard<std::mutex> lk(g_host_mu);
    auto it = g_host_blocks.upper_bound((uintptr_t)p);
    if (it == g_host_blocks.begin()) return nullptr;
    --it;
    const uintptr_t base = it->first;
    if ((uintptr_t)p + len > base + it->second.bytes) return nullptr;
    return it->second.dev + ((uintptr_t)p - base);
}

// ---- pinned arena: a ring of variable-size blocks, freed in any order and
// reclaimed in allocation order --------------------------------------------
struct Arena {
    uint8_t* host = nullptr;
    const uint8_t* dev = nullptr;
    size_t cap = 0, head = 0;
    struct Blk {
        size_t off, bytes;
        bool live;
    };
    std::deque<Blk> blocks;
    uint64_t front_seq = 0; // sequence number of blocks.front()

    bool on_device = false;
    hipError_t init(size_t bytes)
    {
        cap = std::max<size_t>(bytes, 4096) & ~(size_t)15;
        hipError_t e = hipHostMalloc((void**)&host, cap, kDeviceReadFlags);
        if (e != hipSuccess) {
            host = nullptr;
            return e;
        }
        void* d = nullptr;
        e = hipHostGetDevicePointer(&d, host, 0);
        dev = (const uint8_t*)d;
        return e;
    }
    // the same ring over device memory (host is null)
    hipError_t init_device(size_t bytes)
    {
        cap = std::max<size_t>(bytes, 4096) & ~(size_t)15;
        void* d = nullptr;
        const hipError_t e = hipMalloc(&d, cap);
        dev = e == hipSuccess ? (const uint8_t*)d : nullptr;
        on_device = e == hipSuccess;
        return e;
    }
    void destroy()
    {
        if (host) (void)hipHostFree(host);
        if (on_device && dev) (void)hipFree((void*)dev);
        host = nullptr;
        dev = nullptr;
    }
    // Offset of a new block of `bytes`, or -1 while the space is taken.
    int64_t alloc(size_t bytes, uint64_t* seq)
    {
        bytes = (bytes + 15) & ~(size_t)15;
        if (bytes == 0) bytes = 16;
        size_t off;
        if (blocks.empty()) {
            head = 0;
            if (bytes > cap) return -1;
            off = 0;
        } else {
            const size_t tail = blocks.front().off;
            if (head > tail) { // live region [tail, head): free [head, cap) and [0, tail)
                if (cap - head >= bytes) off = head;
                else if (tail >= bytes) off = 0;
                else return -1;
            } else { // wrapped: free [head, tail)
                if (tail - head >= bytes) off = head;
                else return -1;
            }
        }
        head = off + bytes;
        *seq = front_seq + blocks.size();
        blocks.push_back({off, bytes, true});
        return (int64_t)off;
    }
    void release(uint64_t seq)
    {
        blocks[(size_t)(seq - front_seq)].live = false;
        while (!blocks.empty() && !blocks.front().live) {
            blocks.pop_front();
            ++front_seq;
        }
        if (blocks.empty()) head = 0;
    }
};

// ---- per-alphabet tables -------------------------------------------------
struct AlphabetTables {
    int32_t alphabet = GBSW_DNA;
    int32_t alpha = 0;
    int kmax = 0;
    uint8_t lut[256];
    std::vector<uint8_t> bucket_of_len; // host copy (65536)
    std::vector<int> bucket_k;          // K of each bucket index
    std::vector<int> bucket_full;       // full-occupancy grid of each bucket index
    DeviceBuffer d_bucket_of_len, d_accept;
};

struct Request {
    uint64_t ticket = 0;
    int cls = 0;
    int64_t n = 0;
    const uint8_t* ref_src = nullptr; // device-readable, byte 0 of the request's refs
    const uint8_t* qry_src = nullptr;
    std::vector<uint64_t> ro, qo;     // rebased offsets, n + 1
    int32_t* out[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // summary for whole-request admission
    int64_t residues = 0, cells = 0;
    int32_t max_rows = 0, max_cols = 0;
    int64_t bucket_count[kMaxBuckets] = {0};
    int64_t bucket_cells[kMaxBuckets] = {0};
    int32_t bucket_maxcols[kMaxBuckets] = {0};
    // progress (guarded by the server mutex)
    int64_t next = 0;
    int parts = 0, parts_pending = 0;
    int32_t status = GBSW_OK;
    bool done = false, launched = false;
    bool has_arena = false;
    uint64_t arena_seq = 0;
    bool has_darena = false;   // its residues are in the device staging ring
    uint64_t darena_seq = 0;
    Clock::time_point t_submit, t_launch, t_done;
    std::list<Request*>::iterator done_it;
};

struct ScoreClass {
    gbsw_scoring sc;
    int tables = 0; // 0 DNA, 1 protein
    std::deque<Request*> queue;
};

struct Part {
    Request* req;
    int64_t req_begin, mb_begin, count;
};

struct Slot {
    Exec ex;
    Job job;
    gbsw_scoring sc;
    DeviceBuffer seq, meta, out, flags, keys, keys_sorted, vals, order, sort_temp;
    PinnedBuffer meta_host, out_host, flags_host;
    size_t sort_bytes = 0;
    hipEvent_t ev_begin = nullptr, ev_done = nullptr;
    hipEvent_t ev_up = nullptr; // the uploads enqueued before this micro-batch's launch
    std::vector<Part> parts;
    std::vector<GatherTile> tiles;
    int tables = 0;
    bool busy = false;
    uint64_t launch_seq = 0;
    int32_t launch_rc = GBSW_OK;
    Clock::time_point t_admit, t_launched; // host times of the micro-batch (trace)
};

// GBSW_SERVER_TRACE=<path>: one line per micro-batch (host admit and launch
// times, device time, host completion) for diagnosing the serving loop.
FILE* trace_file()
{
    static FILE* f = []() -> FILE* {
        const char* p = std::getenv("GBSW_SERVER_TRACE");
        return p ? std::fopen(p, "w") : nullptr;
    }();
    return f;
}

bool same_scoring(const gbsw_scoring& a, const gbsw_scoring& b)
{
    if (a.alphabet != b.alphabet || a.gap_open != b.gap_open || a.gap_extend != b.gap_extend) return false;
    if (a.alphabet == GBSW_DNA) return a.match == b.match && a.mismatch == b.mismatch;
    return std::memcmp(a.matrix, b.matrix, sizeof(a.matrix)) == 0;
}

} // namespace

struct gbsw_server {
    int device = 0;
    gbsw_server_config cfg;
    AlphabetTables tabs[2];
    Arena arena;              // pinned: residues of requests given in pageable memory
    Arena darena;             // device: every request's residues, uploaded at submit
    hipStream_t up_stream = nullptr; // the uploads (DMA), in submit order
    std::vector<Slot*> slots;
    std::vector<ScoreClass> classes;
    std::unordered_map<uint64_t, Request*> requests;
    std::list<Request*> done_list;
    std::mutex mu;
    std::condition_variable cv_work, cv_done, cv_space;
    std::thread worker;
    bool stop = false;
    uint64_t next_ticket = 1, launch_counter = 0;
    int busy_slots = 0;
    int64_t pending_requests = 0;
    gbsw_serving_stats stats{};
    int64_t device_bytes = 0;

    int init_tables(int which);
    int init_slot(Slot& s);
    bool has_pending() const { return pending_requests > 0; }
    int oldest_class() const;
    void admit(Slot& s);               // under mu
    void complete_unlaunched(Request* r); // under mu
    int launch(Slot& s);               // without mu
    void finish(Slot& s, bool device_ok); // without mu
    void run();
    // frees a completed request's staging blocks (under mu); true if any
    bool release_staging(Request* r)
    {
        bool any = false;
        if (r->has_arena) {
            arena.release(r->arena_seq);
            r->has_arena = false;
            any = true;
        }
        if (r->has_darena) {
            darena.release(r->darena_seq);
            r->has_darena = false;
            any = true;
        }
        return any;
    }
};

namespace {

int alloc_into(DeviceBuffer& b, size_t bytes, int64_t& total)
{
    const hipError_t e = b.reserve(std::max<size_t>(bytes, 16));
    if (e != hipSuccess)
        return fail(e == hipErrorOutOfMemory ? GBSW_E_MEMORY : GBSW_E_DEVICE,
                    std::string("server allocation: ") + hipGetErrorString(e));
    total += (int64_t)b.bytes;
    return GBSW_OK;
}

} // namespace

int gbsw_server::init_tables(int which)
{
    AlphabetTables& t = tabs[which];
    t.alphabet = which == 0 ? GBSW_DNA : GBSW_PROTEIN;
    uint8_t accept[256];
    if (t.alphabet == GBSW_DNA) {
        std::memset(t.lut, 0xFF, 256);
        const size_t na = strnlen(cfg.dna_alphabet, sizeof(cfg.dna_alphabet));
        int A = 0;
        for (size_t i = 0; i < na; ++i) {
            const uint8_t c = (uint8_t)cfg.dna_alphabet[i];
            if (t.lut[c] == 0xFF) t.lut[c] = (uint8_t)A++;
        }
        if (A == 0 || A > kMaxAlpha) return fail(GBSW_E_ARGUMENT, "dna_alphabet must hold 1..31 distinct bytes");
        t.alpha = A;
        for (int c = 0; c < 256; ++c) accept[c] = t.lut[c] != 0xFF;
    } else {
        t.alpha = build_alphabet(GBSW_PROTEIN, nullptr, nullptr, 0, t.lut);
        for (int c = 0; c < 256; ++c) accept[c] = c < 91;
    }
    t.kmax = max_k_for(t.alpha, byte_profile_alphabet(t.alpha));
    t.bucket_of_len = bucket_table(t.alpha, slots[0]->ex.kmodel);
    int nk = 0;
    const int* ks = kernel_k_list(&nk);
    const int nb = bucket_ids();
    if (nb > kMaxBuckets) return fail(GBSW_E_INTERNAL, "too many kernel geometries");
    t.bucket_k.resize((size_t)nb);
    t.bucket_full.resize((size_t)nb);
    for (int i = 0; i < nb; ++i) {
        t.bucket_k[(size_t)i] = ks[i % nk];
        t.bucket_full[(size_t)i] = batched_grid_cap(slots[0]->ex, ks[i % nk], t.alpha);
    }
    int rc = alloc_into(t.d_bucket_of_len, 65536, device_bytes);
    if (!rc) rc = alloc_into(t.d_accept, 256, device_bytes);
    if (rc) return rc;
    GBSW_HIP_TRY(hipMemcpy(t.d_bucket_of_len.ptr, t.bucket_of_len.data(), 65536, hipMemcpyHostToDevice));
    GBSW_HIP_TRY(hipMemcpy(t.d_accept.ptr, accept, 256, hipMemcpyHostToDevice));
    return GBSW_OK;
}

int gbsw_server::init_slot(Slot& s)
{
    const size_t P = (size_t)cfg.pairs_per_slot;
    const size_t R = (size_t)cfg.residues_per_slot;
    const size_t max_tiles = R / kGatherTileBytes + 4 * P + 4;
    int rc = sort_scratch_bytes((int64_t)P, &s.sort_bytes);
    const size_t meta_bytes = sizeof(uint64_t) * 2 * (P + 1) + sizeof(GatherTile) * max_tiles;
    if (!rc) rc = alloc_into(s.seq, R + 64, device_bytes);
    if (!rc) rc = alloc_into(s.meta, meta_bytes, device_bytes);
    if (!rc) rc = alloc_into(s.out, sizeof(int32_t) * 5 * P, device_bytes);
    if (!rc) rc = alloc_into(s.flags, sizeof(int32_t) * P, device_bytes);
    if (!rc) rc = alloc_into(s.keys, 4 * P, device_bytes);
    if (!rc) rc = alloc_into(s.keys_sorted, 4 * P, device_bytes);
    if (!rc) rc = alloc_into(s.vals, 4 * P, device_bytes);
    if (!rc) rc = alloc_into(s.order, 4 * P, device_bytes);
    if (!rc) rc = alloc_into(s.sort_temp, s.sort_bytes, device_bytes);
    if (!rc) rc = alloc_into(s.ex.bnd, (size_t)cfg.boundary_bytes_per_slot, device_bytes);
    if (!rc) rc = alloc_into(s.ex.redo, 4 * P, device_bytes);
    if (!rc) rc = alloc_into(s.ex.queue, 4 * kXcds * kMaxBuckets, device_bytes);
    // split-form bookkeeping: keys, strip gates and counters of every pair, the
    // progress words of up to two boundary slots per workgroup of a full grid
    // per bucket, and each bucket's padding to 256 bytes
    const size_t split_words = 5 * P + 2 * (P / kPairsPerWave + kMaxBuckets) +
                               6 * (size_t)kMaxBuckets * 16 * s.ex.num_cu + (size_t)(kXcds + 64) * kMaxBuckets;
    if (!rc) rc = alloc_into(s.ex.split_buf, 4 * split_words, device_bytes);
    s.ex.fixed = true;
    if (rc) return rc;
    GBSW_HIP_TRY(s.meta_host.reserve(meta_bytes));
    GBSW_HIP_TRY(s.out_host.reserve(sizeof(int32_t) * 5 * P));
    GBSW_HIP_TRY(s.flags_host.reserve(sizeof(int32_t) * P));
    GBSW_HIP_TRY(hipEventCreate(&s.ev_begin));
    GBSW_HIP_TRY(hipEventCreate(&s.ev_done));
    GBSW_HIP_TRY(hipEventCreateWithFlags(&s.ev_up, hipEventDisableTiming));
    return GBSW_OK;
}

int gbsw_server::oldest_class() const
{
    int best = -1;
    uint64_t t = UINT64_MAX;
    for (size_t c = 0; c < classes.size(); ++c)
        if (!classes[c].queue.empty() && classes[c].queue.front()->ticket < t) {
            t = classes[c].queue.front()->ticket;
            best = (int)c;
        }
    return best;
}

// Pack pending requests of the oldest scoring class into slot s (caller holds mu).
void gbsw_server::admit(Slot& s)
{
    const int c = oldest_class();
    ScoreClass& C = classes[(size_t)c];
    const AlphabetTables& T = tabs[C.tables];
    const int nb = (int)T.bucket_k.size();
    const int64_t cap_pairs = cfg.pairs_per_slot, cap_res = cfg.residues_per_slot;
    const int64_t cap_bnd = cfg.boundary_bytes_per_slot;
    int64_t cnt[kMaxBuckets] = {0}, cel[kMaxBuckets] = {0};
    int32_t mcol[kMaxBuckets] = {0};
    int64_t pairs = 0, res = 0, quads = 0, cells = 0;
    int32_t max_rows = 0, max_cols = 0;
    auto quads_of = [&](int b, int64_t count) {
        return std::min<int64_t>(T.bucket_full[(size_t)b], (count + kPairsPerWave - 1) / kPairsPerWave);
    };
    auto bnd_need = [&](int64_t q, int32_t rows) { return q * kPairsPerWave * (int64_t)std::max(1, rows) * 4; };
    s.parts.clear();
    s.sc = C.sc;
    s.tables = C.tables;
    bool full = false;
    while (!C.queue.empty() && !full) {
        Request* r = C.queue.front();
        int64_t take = 0;
        if (r->next == 0) {
            // whole request at once when it fits
            int64_t q2 = 0;
            for (int b = 0; b < nb; ++b) q2 += quads_of(b, cnt[b] + r->bucket_count[b]);
            const int32_t rows2 = std::max(max_rows, r->max_rows);
            if (pairs + r->n <= cap_pairs && res + r->residues <= cap_res && bnd_need(q2, rows2) <= cap_bnd) {
                for (int b = 0; b < nb; ++b) {
                    cnt[b] += r->bucket_count[b];
                    cel[b] += r->bucket_cells[b];
                    mcol[b] = std::max(mcol[b], r->bucket_maxcols[b]);
                }
                quads = q2;
                max_rows = rows2;
                max_cols = std::max(max_cols, r->max_cols);
                res += r->residues;
                cells += r->cells;
                take = r->n;
            }
        }
        if (take == 0) {
            // pair by pair (a split request, or one that does not fit whole)
            for (int64_t p = r->next; p < r->n; ++p) {
                const int32_t la = (int32_t)(r->ro[(size_t)p + 1] - r->ro[(size_t)p]);
                const int32_t lb = (int32_t)(r->qo[(size_t)p + 1] - r->qo[(size_t)p]);
                const int b = T.bucket_of_len[(size_t)std::min(la, lb)];
                const int64_t q2 = quads - quads_of(b, cnt[b]) + quads_of(b, cnt[b] + 1);
                const int32_t rows2 = std::max(max_rows, std::max(la, lb));
                if (pairs + take + 1 > cap_pairs || res + la + lb > cap_res || bnd_need(q2, rows2) > cap_bnd) {
                    full = true;
                    break;
                }
                cnt[b] += 1;
                cel[b] += (int64_t)la * lb;
                mcol[b] = std::max(mcol[b], std::min(la, lb));
                quads = q2;
                max_rows = rows2;
                max_cols = std::max(max_cols, std::min(la, lb));
                res += la + lb;
                cells += (int64_t)la * lb;
                ++take;
            }
        }
        if (take == 0 && pairs == 0) {
            // cannot happen with the capacities checked at creation; never spin on it
            r->status = GBSW_E_INTERNAL;
            r->next = r->n;
            C.queue.pop_front();
            --pending_requests;
            complete_unlaunched(r);
            continue;
        }
        if (take == 0) break;
        s.parts.push_back({r, r->next, pairs, take});
        pairs += take;
        r->next += take;
        r->parts += 1;
        r->parts_pending += 1;
        if (!r->launched) {
            r->launched = true;
            r->t_launch = Clock::now();
        }
        if (r->next == r->n) {
            C.queue.pop_front();
            --pending_requests;
        }
    }
    Job& j = s.job;
    j.alphabet = T.alphabet;
    j.n = pairs;
    j.alpha = T.alpha;
    j.kmodel = s.ex.kmodel;
    j.lut_host = T.lut;
    j.max_rows = max_rows;
    j.max_cols = max_cols;
    j.cells = cells;
    j.buckets.clear();
    int64_t at = 0;
    for (int b = 0; b < nb; ++b) {
        if (!cnt[b]) continue;
        Bucket bk;
        bk.K = T.bucket_k[(size_t)b];
        bk.begin = at;
        bk.count = cnt[b];
        bk.cells = cel[b];
        bk.max_cols = mcol[b];
        at += cnt[b];
        j.buckets.push_back(bk);
    }
    s.busy = true;
    s.launch_seq = ++launch_counter;
    ++busy_slots;
}

// Report a request that no micro-batch holds (caller holds mu).
void gbsw_server::complete_unlaunched(Request* r)
{
    if (r->parts_pending > 0) return; // its earlier parts report it
    r->done = true;
    r->t_done = Clock::now();
    if (release_staging(r)) cv_space.notify_all();
    stats.completed += 1;
    done_list.push_back(r);
    r->done_it = std::prev(done_list.end());
    cv_done.notify_all();
}

// Build the micro-batch's offsets and gather tiles and enqueue its work.
int gbsw_server::launch(Slot& s)
{
    Job& j = s.job;
    const int64_t n = j.n;
    const AlphabetTables& T = tabs[s.tables];
    uint64_t* ro = s.meta_host.as<uint64_t>();
    uint64_t* qo = ro + (n + 1);
    s.tiles.clear();
    // refs packed from 0, queries after the refs (16-byte aligned)
    uint64_t rtot = 0;
    for (const Part& p : s.parts) rtot += p.req->ro[(size_t)(p.req_begin + p.count)] - p.req->ro[(size_t)p.req_begin];
    const uint64_t qbase = (rtot + 15) & ~(uint64_t)15;
    uint64_t rc_at = 0, qc_at = qbase;
    uint8_t* seq = s.seq.as<uint8_t>();
    auto add_tiles = [&](const uint8_t* src, uint64_t dst, uint64_t len, int32_t part) {
        for (uint64_t o = 0; o < len; o += kGatherTileBytes) {
            GatherTile t;
            t.src = src + o;
            t.dst = seq + dst + o;
            t.len = (uint32_t)std::min<uint64_t>(kGatherTileBytes, len - o);
            t.part = part;
            s.tiles.push_back(t);
        }
    };
    for (size_t k = 0; k < s.parts.size(); ++k) {
        const Part& p = s.parts[k];
        const Request* r = p.req;
        const uint64_t r0 = r->ro[(size_t)p.req_begin], q0 = r->qo[(size_t)p.req_begin];
        for (int64_t i = 0; i < p.count; ++i) {
            ro[p.mb_begin + i] = rc_at + (r->ro[(size_t)(p.req_begin + i)] - r0);
            qo[p.mb_begin + i] = qc_at + (r->qo[(size_t)(p.req_begin + i)] - q0);
        }
        const uint64_t rl = r->ro[(size_t)(p.req_begin + p.count)] - r0;
        const uint64_t ql = r->qo[(size_t)(p.req_begin + p.count)] - q0;
        add_tiles(r->ref_src + r0, rc_at, rl, (int32_t)k);
        add_tiles(r->qry_src + q0, qc_at, ql, (int32_t)k);
        rc_at += rl;
        qc_at += ql;
    }
    ro[n] = rc_at;
    qo[n] = qc_at;
    if (qc_at > s.seq.bytes) return fail(GBSW_E_INTERNAL, "micro-batch exceeds its sequence buffer");
    const size_t off_bytes = sizeof(uint64_t) * 2 * (size_t)(n + 1);
    const size_t tiles_at = (off_bytes + 15) & ~(size_t)15;
    const size_t meta_bytes = tiles_at + sizeof(GatherTile) * s.tiles.size();
    if (meta_bytes > s.meta.bytes || meta_bytes > s.meta_host.bytes)
        return fail(GBSW_E_INTERNAL, "micro-batch metadata exceeds its buffer");
    std::memcpy(s.meta_host.as<uint8_t>() + tiles_at, s.tiles.data(), sizeof(GatherTile) * s.tiles.size());

    hipStream_t st = s.ex.stream;
    // the residues of every request of the micro-batch were enqueued on the
    // upload stream before the request was queued: wait for what it holds now
    GBSW_HIP_TRY(hipEventRecord(s.ev_up, up_stream));
    GBSW_HIP_TRY(hipStreamWaitEvent(st, s.ev_up, 0));
    GBSW_HIP_TRY(hipEventRecord(s.ev_begin, st));
    GBSW_HIP_TRY(hipMemcpyAsync(s.meta.ptr, s.meta_host.ptr, meta_bytes, hipMemcpyHostToDevice, st));
    GBSW_HIP_TRY(zero_u32(s.flags.as<uint32_t>(), (int)s.parts.size(), st));
    GBSW_HIP_TRY(gather((const GatherTile*)(s.meta.as<uint8_t>() + tiles_at), (int)s.tiles.size(),
                        T.d_accept.as<uint8_t>(), s.flags.as<int32_t>(), st));
    j.ref = seq;
    j.qry = seq;
    j.ref_off = s.meta.as<uint64_t>();
    j.qry_off = s.meta.as<uint64_t>() + (n + 1);
    j.out = s.out.as<int32_t>();
    j.bucket_of_len = T.d_bucket_of_len.as<uint8_t>();
    j.keys = s.keys.as<uint32_t>();
    j.keys_sorted = s.keys_sorted.as<uint32_t>();
    j.vals = s.vals.as<int32_t>();
    j.order = s.order.as<int32_t>();
    j.sort_temp = s.sort_temp.ptr;
    j.sort_temp_bytes = s.sort_bytes;
    int rc = enqueue_job(s.ex, j, &s.sc, GBSW_BATCHED);
    if (rc) return rc;
    GBSW_HIP_TRY(hipMemcpyAsync(s.out_host.ptr, s.out.ptr, sizeof(int32_t) * 5 * (size_t)n, hipMemcpyDeviceToHost, st));
    GBSW_HIP_TRY(hipMemcpyAsync(s.flags_host.ptr, s.flags.ptr, sizeof(int32_t) * s.parts.size(),
                                hipMemcpyDeviceToHost, st));
    GBSW_HIP_TRY(hipEventRecord(s.ev_done, st));
    return GBSW_OK;
}

// Results of a completed (or failed) micro-batch into its requests.
void gbsw_server::finish(Slot& s, bool device_ok)
{
    const Clock::time_point t_seen = Clock::now();
    int32_t rc = s.launch_rc;
    double busy = 0.0;
    if (rc == GBSW_OK && device_ok) {
        double ms = 0.0;
        rc = complete_job(s.ex, s.job, &s.sc, &ms);
        if (rc == GBSW_OK && s.ex.late_results) {
            // an int32 recompute or an unsplit rerun ran after the copy-back:
            // copy the results again
            hipError_t e = hipMemcpyAsync(s.out_host.ptr, s.out.ptr, sizeof(int32_t) * 5 * (size_t)s.job.n,
                                          hipMemcpyDeviceToHost, s.ex.stream);
            if (e == hipSuccess) e = hipStreamSynchronize(s.ex.stream);
            if (e != hipSuccess) rc = fail(GBSW_E_DEVICE, std::string("result copy: ") + hipGetErrorString(e));
        }
        float t = 0.f;
        if (hipEventElapsedTime(&t, s.ev_begin, s.ev_done) == hipSuccess) busy = t;
    } else if (rc == GBSW_OK) {
        rc = GBSW_E_DEVICE;
    }
    if (rc != GBSW_OK) (void)hipStreamSynchronize(s.ex.stream);
    const int64_t n = s.job.n;
    const int32_t* oh = s.out_host.as<int32_t>();
    const int32_t* fl = s.flags_host.as<int32_t>();
    std::vector<int32_t> part_status(s.parts.size(), GBSW_OK);
    for (size_t k = 0; k < s.parts.size(); ++k) {
        const Part& p = s.parts[k];
        int32_t st = rc;
        if (st == GBSW_OK && fl[k]) st = GBSW_E_RESIDUE;
        part_status[k] = st;
        for (int f = 0; f < 5; ++f) {
            int32_t* dst = p.req->out[f] + p.req_begin;
            if (st == GBSW_OK) std::memcpy(dst, oh + (size_t)f * (size_t)n + (size_t)p.mb_begin, sizeof(int32_t) * (size_t)p.count);
            else std::memset(dst, 0, sizeof(int32_t) * (size_t)p.count);
        }
    }
    const Clock::time_point now = Clock::now();
    if (FILE* tf = trace_file())
        std::fprintf(tf, "mb %llu slot %p alphabet %d pairs %lld parts %zu cells %lld admit_launch_ms %.3f "
                     "device_ms %.3f seen_after_launch_ms %.3f finish_ms %.3f t_ms %.3f\n",
                     (unsigned long long)s.launch_seq, (void*)&s, (int)s.job.alphabet, (long long)s.job.n,
                     s.parts.size(), (long long)s.job.cells, ms_since(s.t_admit, s.t_launched), busy,
                     ms_since(s.t_launched, t_seen), ms_since(t_seen, now), ms_since(trace_epoch(), now));
    std::lock_guard<std::mutex> lk(mu);
    stats.micro_batches += 1;
    stats.busy_ms += busy;
    bool freed = false;
    for (size_t k = 0; k < s.parts.size(); ++k) {
        Request* r = s.parts[k].req;
        if (part_status[k] != GBSW_OK && r->status == GBSW_OK) r->status = part_status[k];
        r->parts_pending -= 1;
        if (r->next == r->n && r->parts_pending == 0 && !r->done) {
            r->done = true;
            r->t_done = now;
            freed = release_staging(r) || freed;
            stats.completed += 1;
            stats.pairs += r->n;
            stats.cells += r->cells;
            done_list.push_back(r);
            r->done_it = std::prev(done_list.end());
        }
    }
    s.parts.clear();
    s.busy = false;
    --busy_slots;
    cv_done.notify_all();
    if (freed) cv_space.notify_all();
}

void gbsw_server::run()
{
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        bool progress = false;
        if (busy_slots > 0) {
            // completions, oldest launch first
            std::vector<Slot*> order;
            for (Slot* s : slots)
                if (s->busy) order.push_back(s);
            std::sort(order.begin(), order.end(), [](Slot* a, Slot* b) { return a->launch_seq < b->launch_seq; });
            lk.unlock();
            for (Slot* s : order) {
                const hipError_t q = s->launch_rc ? hipSuccess : hipEventQuery(s->ev_done);
                if (q == hipErrorNotReady) continue;
                finish(*s, q == hipSuccess);
                progress = true;
            }
            lk.lock();
        }
        while (has_pending()) {
            Slot* free_slot = nullptr;
            for (Slot* s : slots)
                if (!s->busy) {
                    free_slot = s;
                    break;
                }
            if (!free_slot) break;
            free_slot->t_admit = Clock::now();
            admit(*free_slot);
            lk.unlock();
            free_slot->launch_rc = free_slot->parts.empty() ? GBSW_E_INTERNAL : launch(*free_slot);
            free_slot->t_launched = Clock::now();
            lk.lock();
            progress = true;
        }
        if (progress) continue;
        if (busy_slots == 0 && !has_pending()) {
            if (stop) break;
            cv_work.wait(lk);
        } else {
            cv_work.wait_for(lk, std::chrono::microseconds(20));
        }
    }
}

namespace {

void fill_info(const Request* r, gbsw_request_info* info)
{
    if (!info) return;
    info->ticket = r->ticket;
    info->status = r->status;
    info->parts = r->parts;
    info->pairs = r->n;
    info->cells = r->cells;
    info->queue_ms = r->launched ? ms_since(r->t_submit, r->t_launch) : 0.0;
    info->latency_ms = ms_since(r->t_submit, r->t_done);
}

} // namespace

extern "C" {

void gbsw_server_config_default(gbsw_server_config* cfg)
{
    if (!cfg) return;
    std::memset(cfg, 0, sizeof(*cfg));
    cfg->slots = 2;
    cfg->pairs_per_slot = 262144;
    cfg->residues_per_slot = (int64_t)256 << 20;
    cfg->boundary_bytes_per_slot = (int64_t)1 << 30;
    cfg->staging_bytes = (int64_t)256 << 20;
    std::strcpy(cfg->dna_alphabet, "ACGTN");
    cfg->device_staging_bytes = (int64_t)4 << 30;
}

int gbsw_server_destroy(gbsw_server* srv);

int gbsw_server_create(int32_t device, const gbsw_server_config* cfg_in, gbsw_server** out)
{
    if (!out) return fail(GBSW_E_ARGUMENT, "out is null");
    *out = nullptr;
    gbsw_server_config cfg;
    if (cfg_in) cfg = *cfg_in;
    else gbsw_server_config_default(&cfg);
    if (cfg.slots < 1 || cfg.slots > 16) return fail(GBSW_E_ARGUMENT, "slots must be 1..16");
    if (cfg.pairs_per_slot < 1 || cfg.pairs_per_slot > (1 << 26)) return fail(GBSW_E_ARGUMENT, "pairs_per_slot out of range");
    if (cfg.residues_per_slot < (int64_t)(2 * kMaxLen) + 64)
        return fail(GBSW_E_ARGUMENT, "residues_per_slot must hold one pair of maximal length");
    if (cfg.boundary_bytes_per_slot < (int64_t)kPairsPerWave * (int64_t)kMaxLen * 4)
        return fail(GBSW_E_ARGUMENT, "boundary_bytes_per_slot must hold one workgroup at maximal length");
    if (cfg.staging_bytes < 4096) return fail(GBSW_E_ARGUMENT, "staging_bytes must be at least 4096");
    if (cfg.device_staging_bytes < (int64_t)(2 * kMaxLen) + 64)
        return fail(GBSW_E_ARGUMENT, "device_staging_bytes must hold one pair of maximal length");
    int count = 0;
    GBSW_HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(GBSW_E_ARGUMENT, "no such device");
    GBSW_HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    GBSW_HIP_TRY(hipGetDeviceProperties(&prop, device));
    gbsw_server* srv = new gbsw_server();
    srv->device = device;
    srv->cfg = cfg;
    int rc = GBSW_OK;
    for (int i = 0; i < cfg.slots && !rc; ++i) {
        Slot* s = new Slot();
        srv->slots.push_back(s);
        rc = s->ex.init(device, prop.multiProcessorCount);
        if (!rc) rc = srv->init_slot(*s);
    }
    if (!rc) rc = srv->init_tables(0);
    if (!rc) rc = srv->init_tables(1);
    if (!rc) {
        const hipError_t e = srv->arena.init((size_t)cfg.staging_bytes);
        if (e != hipSuccess) rc = fail(GBSW_E_MEMORY, std::string("pinned arena: ") + hipGetErrorString(e));
    }
    if (!rc) {
        hipError_t e = srv->darena.init_device((size_t)cfg.device_staging_bytes);
        if (e == hipSuccess) srv->device_bytes += (int64_t)srv->darena.cap;
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&srv->up_stream, hipStreamNonBlocking);
        if (e != hipSuccess) rc = fail(GBSW_E_MEMORY, std::string("device staging: ") + hipGetErrorString(e));
    }
    if (rc) {
        const std::string err = gbsw_last_error();
        gbsw_server_destroy(srv);
        set_error(err);
        return rc;
    }
    srv->stats.device_bytes = srv->device_bytes;
    srv->worker = std::thread([srv] { srv->run(); });
    *out = srv;
    return GBSW_OK;
}

int gbsw_server_destroy(gbsw_server* srv)
{
    if (!srv) return GBSW_OK;
    if (srv->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->stop = true;
        }
        srv->cv_work.notify_all();
        srv->worker.join();
    }
    (void)hipSetDevice(srv->device);
    if (srv->up_stream) {
        (void)hipStreamSynchronize(srv->up_stream);
        (void)hipStreamDestroy(srv->up_stream);
    }
    for (auto& kv : srv->requests) delete kv.second;
    for (Slot* s : srv->slots) {
        if (s->ev_begin) (void)hipEventDestroy(s->ev_begin);
        if (s->ev_up) (void)hipEventDestroy(s->ev_up);
        if (s->ev_done) (void)hipEventDestroy(s->ev_done);
        s->ex.destroy();
        delete s;
    }
    srv->arena.destroy();
    srv->darena.destroy();
    delete srv;
    return GBSW_OK;
}

int gbsw_server_submit(gbsw_server* srv, const gbsw_scoring* sc, int64_t n, const char* ref, const uint64_t* ref_off,
                       const char* query, const uint64_t* query_off, int32_t* score, int32_t* ref_begin,
                       int32_t* ref_end, int32_t* query_begin, int32_t* query_end, uint64_t* ticket)
{
    if (!srv || !sc || !ticket || n < 0) return fail(GBSW_E_ARGUMENT, "null argument or negative pair count");
    if (sc->alphabet != GBSW_DNA && sc->alphabet != GBSW_PROTEIN) return fail(GBSW_E_ARGUMENT, "unknown alphabet");
    if (n > 0 && (!ref_off || !query_off || !score || !ref_begin || !ref_end || !query_begin || !query_end))
        return fail(GBSW_E_ARGUMENT, "null offsets or result array");
    if (n > (int64_t)0x7FFFFFFF) return fail(GBSW_E_RANGE, "more than 2^31-1 pairs in one request");
    const int which = sc->alphabet == GBSW_DNA ? 0 : 1;
    const AlphabetTables& T = srv->tabs[which];
    Request* r = new Request();
    r->n = n;
    r->out[0] = score;
    r->out[1] = ref_begin;
    r->out[2] = ref_end;
    r->out[3] = query_begin;
    r->out[4] = query_end;
    r->ro.resize((size_t)n + 1);
    r->qo.resize((size_t)n + 1);
    for (int64_t p = 0; p <= n; ++p) {
        r->ro[(size_t)p] = n ? ref_off[p] - ref_off[0] : 0;
        r->qo[(size_t)p] = n ? query_off[p] - query_off[0] : 0;
    }
    for (int64_t p = 0; p < n; ++p) {
        if (r->ro[(size_t)p + 1] < r->ro[(size_t)p] || r->qo[(size_t)p + 1] < r->qo[(size_t)p]) {
            delete r;
            return fail(GBSW_E_ARGUMENT, "offsets must be non-decreasing");
        }
        const uint64_t la = r->ro[(size_t)p + 1] - r->ro[(size_t)p], lb = r->qo[(size_t)p + 1] - r->qo[(size_t)p];
        if (la > kMaxLen || lb > kMaxLen) {
            delete r;
            return fail(GBSW_E_RANGE, "sequences longer than 65000 residues are not supported");
        }
        const int b = T.bucket_of_len[(size_t)std::min(la, lb)];
        r->bucket_count[b] += 1;
        r->bucket_cells[b] += (int64_t)(la * lb);
        r->bucket_maxcols[b] = std::max<int32_t>(r->bucket_maxcols[b], (int32_t)std::min(la, lb));
        r->cells += (int64_t)(la * lb);
        r->max_rows = std::max<int32_t>(r->max_rows, (int32_t)std::max(la, lb));
        r->max_cols = std::max<int32_t>(r->max_cols, (int32_t)std::min(la, lb));
    }
    const uint64_t rbytes = r->ro[(size_t)n], qbytes = r->qo[(size_t)n];
    r->residues = (int64_t)(rbytes + qbytes);
    if ((rbytes && !ref) || (qbytes && !query)) {
        delete r;
        return fail(GBSW_E_ARGUMENT, "sequence buffer is null");
    }
    {
        // scoring rules and the score range, as for a batch of these pairs
        Job probe;
        probe.alphabet = T.alphabet;
        probe.alpha = T.alpha;
        probe.max_cols = r->max_cols;
        std::vector<int16_t> sub;
        int max_sub = 0;
        bool p8 = false;
        const int rc = prepare_scoring(probe, sc, sub, &max_sub, &p8);
        if (rc) {
            delete r;
            return rc;
        }
    }
    // Residues of a request of at least kDmaMinBytes go to the device staging
    // ring now, by DMA on the upload stream (one copy when ref and query bytes
    // are adjacent), so that they are in HBM while the request waits for a slot
    // and its micro-batch's gather reads HBM.  Smaller requests, for which a
    // DMA's fixed cost (about 12 us per copy) would dominate, are read by the
    // gather kernel from pinned host memory.  Pinned sources (gbsw_host_alloc)
    // are read in place; pageable ones are copied into the pinned arena first.
    const uint8_t* rp = n ? (const uint8_t*)ref + ref_off[0] : nullptr;
    const uint8_t* qp = n ? (const uint8_t*)query + query_off[0] : nullptr;
    const uint8_t* rdev = rbytes ? registered_device_ptr(rp, rbytes) : nullptr;
    const uint8_t* qdev = qbytes ? registered_device_ptr(qp, qbytes) : nullptr;
    const uint64_t copy_r = (rbytes && !rdev) ? rbytes : 0;
    const uint64_t copy_q = (qbytes && !qdev) ? qbytes : 0;
    const bool dma = rbytes + qbytes >= kDmaMinBytes;
    if (rbytes + qbytes) {
        // pinned arena block: the ref bytes, then the query bytes right after them
        const size_t pneed = (size_t)(copy_r + copy_q);
        const size_t dneed = (size_t)(rbytes + qbytes);
        std::unique_lock<std::mutex> lk(srv->mu);
        if (dma && ((dneed + 15) & ~(size_t)15) > srv->darena.cap) {
            delete r;
            return fail(GBSW_E_MEMORY, "request larger than the device staging ring (device_staging_bytes)");
        }
        if (pneed && ((pneed + 15) & ~(size_t)15) > srv->arena.cap) {
            delete r;
            return fail(GBSW_E_MEMORY, "request larger than the staging arena (use gbsw_host_alloc buffers)");
        }
        int64_t doff = 0, poff = 0;
        if (dma) {
            uint64_t dseq = 0;
            while ((doff = srv->darena.alloc(dneed, &dseq)) < 0) srv->cv_space.wait(lk);
            r->has_darena = true;
            r->darena_seq = dseq;
        }
        if (pneed) {
            uint64_t pseq = 0;
            while ((poff = srv->arena.alloc(pneed, &pseq)) < 0) srv->cv_space.wait(lk);
            r->has_arena = true;
            r->arena_seq = pseq;
        }
        lk.unlock();
        // host (and device-mapped) addresses of the bytes to move
        const uint8_t* hr = rp;
        const uint8_t* hq = qp;
        const uint8_t* mr = rdev;
        const uint8_t* mq = qdev;
        if (copy_r) {
            std::memcpy(srv->arena.host + poff, rp, copy_r);
            hr = srv->arena.host + poff;
            mr = srv->arena.dev + poff;
        }
        if (copy_q) {
            std::memcpy(srv->arena.host + poff + copy_r, qp, copy_q);
            hq = srv->arena.host + poff + copy_r;
            mq = srv->arena.dev + poff + copy_r;
        }
        if (!dma) {
            r->ref_src = mr;
            r->qry_src = mq;
        } else {
            uint8_t* dd = (uint8_t*)srv->darena.dev + doff;
            hipError_t e = hipSuccess;
            if (rbytes && qbytes && hq == hr + rbytes) {
                e = hipMemcpyAsync(dd, hr, rbytes + qbytes, hipMemcpyHostToDevice, srv->up_stream);
            } else {
                if (rbytes) e = hipMemcpyAsync(dd, hr, rbytes, hipMemcpyHostToDevice, srv->up_stream);
                if (e == hipSuccess && qbytes)
                    e = hipMemcpyAsync(dd + rbytes, hq, qbytes, hipMemcpyHostToDevice, srv->up_stream);
            }
            if (e != hipSuccess) {
                (void)hipStreamSynchronize(srv->up_stream); // nothing of this request stays in flight
                std::lock_guard<std::mutex> lk2(srv->mu);
                if (srv->release_staging(r)) srv->cv_space.notify_all();
                delete r;
                return fail(GBSW_E_DEVICE, std::string("request upload: ") + hipGetErrorString(e));
            }
            r->ref_src = dd;
            r->qry_src = dd + rbytes;
        }
    }
    std::lock_guard<std::mutex> lk(srv->mu);
    r->ticket = srv->next_ticket++;
    r->t_submit = Clock::now();
    int cls = -1;
    for (size_t c = 0; c < srv->classes.size(); ++c)
        if (same_scoring(srv->classes[c].sc, *sc)) {
            cls = (int)c;
            break;
        }
    if (cls < 0) {
        ScoreClass C;
        C.sc = *sc;
        C.tables = which;
        srv->classes.push_back(C);
        cls = (int)srv->classes.size() - 1;
    }
    r->cls = cls;
    srv->requests[r->ticket] = r;
    srv->stats.submitted += 1;
    *ticket = r->ticket;
    if (n == 0) {
        r->done = true;
        r->t_done = r->t_submit;
        srv->stats.completed += 1;
        srv->done_list.push_back(r);
        r->done_it = std::prev(srv->done_list.end());
        srv->cv_done.notify_all();
        return GBSW_OK;
    }
    srv->classes[(size_t)cls].queue.push_back(r);
    srv->pending_requests += 1;
    srv->cv_work.notify_one();
    return GBSW_OK;
}

int gbsw_server_wait(gbsw_server* srv, uint64_t ticket, double timeout_s, gbsw_request_info* info)
{
    if (!srv) return fail(GBSW_E_ARGUMENT, "server is null");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->requests.find(ticket);
    if (it == srv->requests.end()) return fail(GBSW_E_ARGUMENT, "unknown or already reported ticket");
    Request* r = it->second;
    if (timeout_s < 0) {
        srv->cv_done.wait(lk, [r] { return r->done; });
    } else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [r] { return r->done; })) {
        return fail(GBSW_E_TIMEOUT, "request not complete");
    }
    fill_info(r, info);
    srv->done_list.erase(r->done_it);
    srv->requests.erase(it);
    delete r;
    return GBSW_OK;
}

int gbsw_server_poll(gbsw_server* srv, int32_t max, gbsw_request_info* infos, int32_t* count)
{
    if (!srv || !count || max < 0 || (max > 0 && !infos)) return fail(GBSW_E_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    int32_t k = 0;
    while (k < max && !srv->done_list.empty()) {
        Request* r = srv->done_list.front();
        srv->done_list.pop_front();
        fill_info(r, &infos[k++]);
        srv->requests.erase(r->ticket);
        delete r;
    }
    *count = k;
    return GBSW_OK;
}

int gbsw_server_stats(gbsw_server* srv, gbsw_serving_stats* out)
{
    if (!srv || !out) return fail(GBSW_E_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(srv->mu);
    *out = srv->stats;
    out->pending = srv->pending_requests;
    out->device_bytes = srv->device_bytes;
    return GBSW_OK;
}

int gbsw_host_alloc(uint64_t bytes, void** out)
{
    if (!out || bytes == 0) return fail(GBSW_E_ARGUMENT, "null argument or zero size");
    *out = nullptr;
    void* p = nullptr;
    hipError_t e = hipHostMalloc(&p, (size_t)bytes, kDeviceReadFlags);
    if (e != hipSuccess) return fail(GBSW_E_MEMORY, std::string("hipHostMalloc: ") + hipGetErrorString(e));
    void* d = nullptr;
    e = hipHostGetDevicePointer(&d, p, 0);
    if (e != hipSuccess) {
        (void)hipHostFree(p);
        return fail(GBSW_E_DEVICE, std::string("hipHostGetDevicePointer: ") + hipGetErrorString(e));
    }
    std::lock_guard<std::mutex> lk(g_host_mu);
    g_host_blocks[(uintptr_t)p] = HostBlock{(size_t)bytes, (const uint8_t*)d};
    *out = p;
    return GBSW_OK;
}

int gbsw_host_free(void* p)
{
    if (!p) return GBSW_OK;
    {
        std::lock_guard<std::mutex> lk(g_host_mu);
        auto it = g_host_blocks.find((uintptr_t)p);
        if (it == g_host_blocks.end()) return fail(GBSW_E_ARGUMENT, "not a gbsw_host_alloc block");
        g_host_blocks.erase(it);
    }
    GBSW_HIP_TRY(hipHostFree(p));
    return GBSW_OK;
}

} // extern "C"
