/*
 * gbsw_oracle.h -- CPU restatement of GPU-BSW (ADEPT) batched Smith-Waterman.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP engine is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path (libgbsw.so) never links,
 * loads or calls it.
 *
 * Every routine restates the reference algorithm with int32 arithmetic:
 *   forward  : kernel.cpp:198-394 (DNA), 605-828 (protein)
 *   reverse  : kernel.cpp:397-604 (DNA), 830-1059 (protein)
 *   tie rules: kernel.cpp:47-86 + 137-182 (forward), 3-45 + 90-135 (reverse)
 *   encoding : driver.cpp:252-258
 */
#ifndef GBSW_ORACLE_H
#define GBSW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { GBSW_ORACLE_DNA = 0, GBSW_ORACLE_AA = 1 };

typedef struct {
    int32_t kind;        /* GBSW_ORACLE_DNA or GBSW_ORACLE_AA */
    int32_t match;       /* DNA: score when the two bytes are equal        */
    int32_t mismatch;    /* DNA: score when they differ                    */
    int32_t gap_open;    /* score of the first position of a gap (<= 0)    */
    int32_t gap_extend;  /* score of every further gap position (<= 0)     */
    int16_t matrix[576]; /* AA: 24x24, index enc(row)*24 + enc(col)        */
} gbsw_oracle_scoring;

typedef struct {
    int32_t score;
    int32_t ref_begin;   /* 0-based */
    int32_t ref_end;     /* 1-based (exclusive 0-based), as the reference */
    int32_t query_begin; /* 0-based */
    int32_t query_end;   /* 1-based */
} gbsw_oracle_result;

/* Reference residue encoding (driver.cpp:252-258): 0..23, -1 for bytes the
 * reference cannot index (>= 91 or negative). */
int gbsw_oracle_encode_aa(int byte);

/* Tie-break priority of column t of the shorter sequence in the forward
 * arg-max (kernel.cpp:47-86, 137-182): lower wins. */
uint32_t gbsw_oracle_prio(uint32_t t);

/* Forward pass of one pair: score and 1-based ends. Returns 0, or -1 on an
 * invalid residue. */
int gbsw_oracle_forward(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                        const gbsw_oracle_scoring* sc, int32_t* score, int32_t* ref_end,
                        int32_t* query_end);

/* Reverse pass of one pair on the prefixes [0,ref_end) x [0,query_end).
 * Writes 0-based begins and the reverse-DP maximum. */
int gbsw_oracle_reverse(const char* ref, int32_t ref_end, const char* query, int32_t query_end,
                        const gbsw_oracle_scoring* sc, int32_t* rev_max, int32_t* ref_begin,
                        int32_t* query_begin);

/* Full alignment of one pair (forward then reverse). Score-0 pairs return all
 * five fields 0 (the reference's behaviour there is undefined). */
int gbsw_oracle_align(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                      const gbsw_oracle_scoring* sc, gbsw_oracle_result* out);

/* Batch driver.  Pair k is ref[ref_off[k] .. ref_off[k+1]) against
 * query[query_off[k] .. query_off[k+1]).  nthreads <= 0 uses every core.
 * Returns 0, or -(k+1) for the first pair with an invalid residue. */
int64_t gbsw_oracle_align_batch(int64_t n, const char* ref, const uint64_t* ref_off,
                                const char* query, const uint64_t* query_off,
                                const gbsw_oracle_scoring* sc, gbsw_oracle_result* out,
                                int nthreads);

/* Alignment path of one aligned pair (no counterpart in the reference, which
 * only declares traceBack, kernel.hpp:36-39): the best global alignment of
 * ref[res->ref_begin, res->ref_end) against query[res->query_begin,
 * res->query_end) under the forward pass's scoring and orientation; among
 * optimal ones, the op sequence read from its end is lexicographically smallest
 * with M < D < I.  ops receive BAM CIGAR words (length << 4 | op; op 0 = M,
 * 1 = I, 2 = D), at most cap; *path_score its score.  Pairs with score 0 give no
 * ops.  Returns 0, -1 on an invalid residue, -3 when cap is too small. */
int gbsw_oracle_traceback(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                          const gbsw_oracle_scoring* sc, const gbsw_oracle_result* res, uint32_t* ops,
                          int32_t cap, int32_t* nops, int32_t* path_score);

/* Batch form: pair k writes at most cap_off[k+1] - cap_off[k] words at
 * ops + cap_off[k], their count to nops[k].  Returns 0 or -(k+1) for the first
 * failing pair. */
int64_t gbsw_oracle_traceback_batch(int64_t n, const char* ref, const uint64_t* ref_off, const char* query,
                                    const uint64_t* query_off, const gbsw_oracle_scoring* sc,
                                    const gbsw_oracle_result* res, uint32_t* ops, const int64_t* cap_off,
                                    int32_t* nops, int32_t* path_score, int nthreads);

/* Threads an OpenMP region actually gets for nthreads (<= 0: all). */
int gbsw_oracle_threads(int nthreads);

#ifdef __cplusplus
}
#endif
#endif
