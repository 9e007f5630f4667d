/*
 * gbsw.h -- C ABI of the MI355X batched Smith-Waterman engine (libgbsw.so).
 *
 * This is the boundary a host program binds to (ctypes, cgo, JNI, N-API ...):
 * plain pointers and sizes, status codes, no C++ or framework types.  The C++
 * drop-in gpu_bsw_driver::kernel_driver_dna / kernel_driver_aa
 * (include/gpu_bsw_driver.hpp) is built on these calls.
 *
 * Semantics follow the reference exactly (GPU-BSW / ADEPT):
 *   pair k aligns ref k (the reference's "contig", seqA) with query k (the
 *   reference's "read", seqB), driver.cpp:16-21 / 246-250;
 *   outputs per pair: top score, 0-based begins, 1-based ends, with the
 *   reference's orientation and tie rules (kernel.cpp:198-1059);
 *   DNA scores compare raw bytes (kernel.cpp:366); protein residues are encoded
 *   with the reference's 91-entry table (driver.cpp:252-258) and scored with a
 *   caller-supplied 24x24 matrix indexed enc(row)*24 + enc(col) (kernel.cpp:785-788).
 * Differences, by design: results are int32; errors are returned as status
 * codes instead of exit() (utils.hpp:5-21); one device per context.
 */
#ifndef GBSW_H
#define GBSW_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define GBSW_OK 0
#define GBSW_E_ARGUMENT (-1)  /* null pointer, bad size, bad mode */
#define GBSW_E_RESIDUE (-2)   /* protein byte >= 91 or negative; DNA alphabet > 31 symbols */
#define GBSW_E_SCORING (-3)   /* gap scores must be <= 0 */
#define GBSW_E_DEVICE (-4)    /* HIP runtime error (message in gbsw_last_error) */
#define GBSW_E_MEMORY (-5)    /* device or pinned allocation failed */
#define GBSW_E_RANGE (-6)     /* scores or lengths beyond the engine's exact range */
#define GBSW_E_INTERNAL (-7)  /* an internal consistency check failed */

#define GBSW_DNA 0
#define GBSW_PROTEIN 1

/* Execution forms.
 * PLAIN:   the reference's shape -- one forward launch and one reverse launch per
 *          batch, pairs in input order, one kernel geometry sized for the longest
 *          pair of the batch (driver.cpp:169-208).
 * BATCHED: pairs binned by length, one persistent kernel per length bucket pulling
 *          work from a device queue, forward and reverse fused, geometry chosen
 *          per bucket. */
#define GBSW_PLAIN 0
#define GBSW_BATCHED 1

typedef struct gbsw_context gbsw_context;
typedef struct gbsw_batch gbsw_batch;

typedef struct {
    int32_t alphabet;      /* GBSW_DNA or GBSW_PROTEIN */
    int16_t match;         /* DNA: score of equal bytes   (reference scores[0]) */
    int16_t mismatch;      /* DNA: score of unequal bytes (scores[1])           */
    int16_t gap_open;      /* first gap position, <= 0    (scores[2] / openGap)  */
    int16_t gap_extend;    /* each further position, <= 0 (scores[3] / extendGap)*/
    int16_t matrix[576];   /* protein: 24x24, ARNDCQEGHILKMFPSTWYVBZX* order     */
} gbsw_scoring;

typedef struct {
    double device_ms;       /* last gbsw_batch_align: device time, hipEvents        */
    double main_kernel_ms;  /* duration of its longest kernel launch                */
    int64_t cells;          /* sum of |ref|*|query| (the reference's "Total Cells") */
    int32_t launches;       /* kernel launches issued                               */
    int32_t buckets;        /* length buckets used (BATCHED)                        */
    int32_t graph;          /* 0 issued call by call, 1 captured into a hipGraph
                               and launched, 2 a replay of the captured graph       */
    int32_t reserved;
} gbsw_stats;

/* Create a context on one HIP device.  A context owns a stream, scratch buffers
 * and events; calls on one context must not overlap. */
int gbsw_context_create(int32_t device, gbsw_context** out);
int gbsw_context_destroy(gbsw_context* ctx);

/* Context options (gbsw_context_set_option).
 * STREAMS: streams the length buckets of the batched form run on, 1..8 (default
 *          4, or the GBSW_STREAMS environment variable); 1 serialises them, so
 *          every launch is timed alone.
 * F16:     1 (default; GBSW_F16=0 turns it off): length buckets whose scores
 *          provably stay below 2048 run an exact integer-valued f16 kernel;
 *          0: every bucket runs the int16 kernel.  Results are identical. */
#define GBSW_OPT_STREAMS 1
#define GBSW_OPT_F16 2
/* GEOMETRY: how a pair's shorter length picks its kernel geometry (K columns per
 *          lane), for batches created afterwards.  0: fewest padded columns and
 *          steps, strips x (K + 2); 1 (default; GBSW_KMODEL=m selects model m):
 *          that cost divided by min(1, waves per SIMD / 2.5), the share of the
 *          VALU issue rate the geometry's occupancy allows; 2..5: the same with
 *          2.0, 3.0, 4.0 or 1.5 waves (gbsw_batch_autotune picks among them).  Applies to the length buckets of
 *          gbsw_batch_align(GBSW_BATCHED), and to the plain form for protein
 *          (the plain form uses 0 for DNA).  Results are identical. */
#define GBSW_OPT_GEOMETRY 3
/* SPLIT:   1 (default; GBSW_SPLIT=0 turns it off): length buckets of pairs that
 *          span several strips, with too few pairs to occupy the GPU twice
 *          over, run each strip as its own work item, on several workgroups of
 *          one XCD in a pipeline (a strip follows the boundary rows of the one
 *          before as they are written); 0: one workgroup runs a pair's strips
 *          one after another.  Results are identical. */
#define GBSW_OPT_SPLIT 4
/* GRAPHS:  1 (default; GBSW_GRAPHS=0 turns it off): gbsw_batch_align in the
 *          batched form captures its stream work (table copies, binning, sort,
 *          every bucket launch on every stream) into a hipGraph the first time,
 *          and replays it while the batch, scoring, options and context buffers
 *          stay the same (gbsw_align, which aligns its batch once, does not
 *          capture); 0: the work is issued call by call.  Per-launch durations
 *          (gbsw_batch_launch_stats) may be 0 for graph launches.  Results are
 *          identical. */
#define GBSW_OPT_GRAPHS 5
int gbsw_context_set_option(gbsw_context* ctx, int32_t option, int64_t value);

/* Message of the last error on this thread ("" if none). */
const char* gbsw_last_error(void);

/* Upload a batch into device memory.  Pair k is
 *   ref[ref_off[k] .. ref_off[k+1])  against  query[query_off[k] .. query_off[k+1])
 * (n+1 offsets each, non-decreasing).  Residues are validated for the alphabet
 * here.  Host buffers may be reused after the call returns. */
int gbsw_batch_create(gbsw_context* ctx, int32_t alphabet, int64_t n, const char* ref,
                      const uint64_t* ref_off, const char* query, const uint64_t* query_off,
                      gbsw_batch** out);
int gbsw_batch_destroy(gbsw_batch* batch);

/* Align every pair of a resident batch; results stay in device memory. */
int gbsw_batch_align(gbsw_batch* batch, const gbsw_scoring* scoring, int32_t mode);

/* Copy the results of the last gbsw_batch_align to host arrays of n int32 each
 * (any pointer may be NULL to skip that field). */
int gbsw_batch_results(gbsw_batch* batch, int32_t* score, int32_t* ref_begin, int32_t* ref_end,
                       int32_t* query_begin, int32_t* query_end);

/* Device pointer of the five result arrays (score, ref_begin, ref_end,
 * query_begin, query_end; n int32 each, back to back). */
int gbsw_batch_device_results(gbsw_batch* batch, const int32_t** d_out);

int gbsw_batch_stats(const gbsw_batch* batch, gbsw_stats* out);

/* Autotune the length -> geometry map on this batch: aligns it (BATCHED,
 * `repeats` timed runs after one untimed) under each distinct map of the
 * GBSW_OPT_GEOMETRY models, keeps the fastest (by device time; another model
 * must beat the batch's current one by more than 1.5 %) for the batch and makes
 * it the context's model for batches created afterwards.  *model and *ms
 * (either may be null) receive the choice and its time.  Results do not depend
 * on the map; the batch's result arrays hold those of the last run. */
int gbsw_batch_autotune(gbsw_batch* batch, const gbsw_scoring* scoring, int32_t repeats, int32_t* model, double* ms);

/* Alignment paths (CIGAR) of the pairs of the last gbsw_batch_align.  No
 * counterpart in the reference, which declares a traceback (traceBack,
 * kernel.hpp:36-39) but never defines it.  Pair k's path is the best global
 * alignment of ref[ref_begin, ref_end) against query[query_begin, query_end)
 * under `scoring` (a run of L gap positions scores gap_open + (L-1) gap_extend;
 * gap_open <= gap_extend is required; substitution s(row, column) with rows the
 * longer full sequence, as in the forward pass); among optimal alignments, the
 * one whose op sequence read from its end is lexicographically smallest with
 * M < D < I.  Ops are BAM CIGAR words: length << 4 | op, op 0 = M (a residue
 * pair, match or mismatch), 1 = I (a query residue against a gap), 2 = D (a ref
 * residue against a gap).  Pairs with score 0 have no ops.
 *   cigar_offsets: n+1 entries, always filled: pair k's words are
 *                  cigar[cigar_offsets[k] .. cigar_offsets[k+1]).
 *   cigar:         capacity words; when cigar_offsets[n] > capacity nothing is
 *                  copied and GBSW_E_RANGE is returned (call again with a larger
 *                  buffer; rows + columns per pair always suffice).
 *   path_score:    n entries or NULL: the score of each path (equal to the pair's
 *                  score whenever its begin and end belong to one best local
 *                  alignment).
 *   device_ms:     NULL or the duration of the path kernels (hipEvents).
 * Device scratch: about one byte per cell of the pairs' rectangles, processed in
 * launches of at most 12 GiB (a third of the free device memory when that is
 * less), two of which are held at once; it stays with the batch until
 * gbsw_batch_destroy. */
int gbsw_batch_traceback(gbsw_batch* batch, const gbsw_scoring* scoring, int64_t capacity, uint32_t* cigar,
                         int64_t* cigar_offsets, int32_t* path_score, double* device_ms);

/* One kernel launch of the last gbsw_batch_align (index 0 .. stats.launches-1). */
typedef struct {
    int32_t kind;             /* 0 plain forward, 1 plain reverse, 2 batched fwd+rev
                                 int16, 3 batched int32 (saturated pairs), 4 batched
                                 f16; +256: scores from the substitution table
                                 (alphabets of 8+ codes) instead of a profile       */
    int32_t columns_per_lane; /* K: columns per virtual lane (strip = 32*K columns)  */
    int32_t workgroups;       /* grid size (64-lane workgroups)                       */
    int64_t pairs;            /* pairs processed by the launch                        */
    int64_t cells;            /* their sum of |ref|*|query| (0 for the reverse launch)*/
    double ms;                /* duration, hipEvents around the launch                */
} gbsw_launch_stat;
int gbsw_batch_launch_stats(const gbsw_batch* batch, int32_t index, gbsw_launch_stat* out);

/* One call, host memory in and out: upload, align, download (the drop-in path).
 * The context keeps the device buffers of this call's batch for later calls
 * (they only grow) and frees them in gbsw_context_destroy. */
int gbsw_align(gbsw_context* ctx, const gbsw_scoring* scoring, int32_t mode, int64_t n,
               const char* ref, const uint64_t* ref_off, const char* query,
               const uint64_t* query_off, int32_t* score, int32_t* ref_begin, int32_t* ref_end,
               int32_t* query_begin, int32_t* query_end);

/* gbsw_align without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device or pinned memory (tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, results and errors as
 * gbsw_align, but every device buffer of the call -- sequences, offsets,
 * results, binning and sort arrays, strip-boundary rows, counters -- is carved
 * from one static device array per device of 256 MiB (set up by
 * gbsw_context_create; each device the library runs on holds it until the
 * process ends), and the call makes no hipMalloc, hipFree or pinned
 * allocation.  A call whose sequences and per-pair arrays (offsets, results,
 * binning and sort arrays: about 60 bytes per pair) leave less than 1 MiB of
 * the array free is
 * rejected with GBSW_E_RANGE before any device work; boundary rows that do not
 * fit the rest run on fewer workgroups.  All contexts of one device share its
 * array: calls that overlap on one device race (serialise them). */
int gbsw_align_noalloc(gbsw_context* ctx, const gbsw_scoring* scoring, int32_t mode, int64_t n,
                       const char* ref, const uint64_t* ref_off, const char* query,
                       const uint64_t* query_off, int32_t* score, int32_t* ref_begin, int32_t* ref_end,
                       int32_t* query_begin, int32_t* query_end);

/* Total device memory in bytes (driver.cpp:10-14, get_tot_gpu_mem). */
int gbsw_device_memory(int32_t device, uint64_t* bytes);

/* ------------------------------------------------------------------------
 * Serving: an asynchronous request loop with continuous batching.
 *
 * No counterpart in the reference, whose driver is one blocking call per batch
 * (driver.hpp:32-37).  A server owns `slots` micro-batch slots, each a set of
 * device-resident buffers (sequences, offsets, results, binning and boundary
 * scratch) of fixed capacity, with its own streams.  A scheduler thread packs
 * pending requests (oldest first, same scoring) into a free slot as soon as one
 * frees up, so requests that arrive while others run are admitted and overlap
 * on the device; admission is bounded by each slot's HBM footprint (pairs,
 * residue bytes, strip-boundary bytes) and large requests are split across
 * micro-batches.  Results are bit-identical to gbsw_align on the same pairs,
 * whatever the mix they were batched with.
 * ------------------------------------------------------------------------ */
#define GBSW_E_TIMEOUT (-8)   /* gbsw_server_wait: the request is not complete yet */

typedef struct gbsw_server gbsw_server;

typedef struct {
    int32_t slots;                   /* micro-batches in flight (default 2)                 */
    int32_t pairs_per_slot;          /* pair capacity of one micro-batch (default 262144)   */
    int64_t residues_per_slot;       /* ref+query bytes of one micro-batch (default 256 MiB) */
    int64_t boundary_bytes_per_slot; /* strip-boundary scratch of one micro-batch (default 1 GiB) */
    int64_t staging_bytes;           /* pinned host arena for copied requests (default 256 MiB) */
    char dna_alphabet[32];           /* DNA bytes accepted, NUL-terminated (default "ACGTN"); other
                                        bytes fail the request with GBSW_E_RESIDUE          */
    int64_t device_staging_bytes;    /* HBM ring holding the residues of submitted, unfinished
                                        requests of 256 KiB of residues or more, uploaded by DMA
                                        at submit (default 4 GiB); submit blocks while it is
                                        full                                                 */
} gbsw_server_config;

typedef struct {
    uint64_t ticket;
    int32_t status;       /* GBSW_OK or the request's error code */
    int32_t parts;        /* micro-batches the request was split over */
    int64_t pairs;
    int64_t cells;
    double queue_ms;      /* submit -> first micro-batch launched */
    double latency_ms;    /* submit -> results written */
} gbsw_request_info;

typedef struct {
    int64_t submitted, completed, pending, micro_batches;
    int64_t pairs, cells;            /* of completed requests                     */
    double busy_ms;                  /* device time summed over micro-batches     */
    int64_t device_bytes;            /* HBM held by the server                    */
} gbsw_serving_stats;

void gbsw_server_config_default(gbsw_server_config* cfg);
int gbsw_server_create(int32_t device, const gbsw_server_config* cfg, gbsw_server** out);
/* Completes every submitted request, then stops the scheduler and frees the server. */
int gbsw_server_destroy(gbsw_server* srv);

/* Queue n pairs (layout as gbsw_batch_create) for alignment under `scoring`.
 * Returns at once with a ticket; the five n-entry int32 result arrays are
 * written before the request is reported complete and must stay valid until
 * then.  Input buffers allocated with gbsw_host_alloc are read in place (by
 * DMA into the server's device staging ring at submit for requests of 256 KiB
 * of residues or more, by the device over PCIe otherwise) and must stay
 * unchanged until completion; other input buffers are first copied into the
 * server's pinned arena and may be reused when this returns.  Thread-safe. */
int gbsw_server_submit(gbsw_server* srv, const gbsw_scoring* scoring, int64_t n, const char* ref,
                       const uint64_t* ref_off, const char* query, const uint64_t* query_off,
                       int32_t* score, int32_t* ref_begin, int32_t* ref_end, int32_t* query_begin,
                       int32_t* query_end, uint64_t* ticket);
/* Block until the request completes (timeout_s < 0: no limit) and report it.
 * A reported ticket is released and unknown afterwards. */
int gbsw_server_wait(gbsw_server* srv, uint64_t ticket, double timeout_s, gbsw_request_info* info);
/* Report up to `max` completed requests not reported yet, in completion order. */
int gbsw_server_poll(gbsw_server* srv, int32_t max, gbsw_request_info* infos, int32_t* count);
int gbsw_server_stats(gbsw_server* srv, gbsw_serving_stats* out);

/* Pinned host memory the device reads in place (zero-copy submission). */
int gbsw_host_alloc(uint64_t bytes, void** out);
int gbsw_host_free(void* p);

/* Version string of the library. */
const char* gbsw_version(void);

#ifdef __cplusplus
}
#endif
#endif
