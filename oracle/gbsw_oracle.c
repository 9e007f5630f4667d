/*
 * gbsw_oracle.c -- scalar CPU restatement of GPU-BSW (ADEPT) local alignment.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  See gbsw_oracle.h.
 *
 * The reference computes, per pair, an affine-gap (Gotoh) local alignment on
 * anti-diagonals with one CUDA thread per column of the shorter sequence.  This
 * file computes the same cells row by row with int32 arithmetic and applies the
 * reference's tie rules in closed form:
 *
 *  Orientation (kernel.cpp:248-259): rows R = the longer sequence, columns C =
 *    the shorter; on equal length R = ref (seqA), C = query (seqB).
 *  Recurrence (kernel.cpp:291-367), 1-based rows i, columns j:
 *    F(i,j) = max(F(i-1,j)+ext, H(i-1,j)+open), F(0,.) = H(0,.) = 0
 *             (zeroed registers, 270-272, 291-307: gives F(1,j)=max(ext,open))
 *    E(i,1) = 0 (lane 0 of warp 0 forces eVal = heVal = 0, 346-350)
 *    E(i,j) = max(E(i,j-1)+ext, H(i,j-1)+open) for j > 1 (331-352)
 *    D      = H(i-1,j-1) + s(R[i-1], C[j-1])   (H(.,0) = 0, 354-366)
 *    H(i,j) = max(D, F, E, 0)                  (findMaxFour, 186-196, 367)
 *  Substitution: DNA raw byte equality (366); protein
 *    M[enc(R[i-1])*24 + enc(C[j-1])] (785-788, encoding driver.cpp:252-258).
 *  Forward result (368-393): score = max H; each column keeps the first row
 *    reaching its max (strict '>' update, 368-370); the arg-max reduction
 *    (47-86, 137-182) prefers the larger value, then the smaller row, then the
 *    lower tree position prio(t).  Hence the winner is the cell with H = score
 *    that minimises (row, prio(column)).
 *  Reverse result (397-604): the same DP on the reversed prefixes
 *    A[0:ref_end), B[0:query_end), orientation decided by the two end lengths
 *    (444-459).  Each column keeps its first reversed row reaching its max; the
 *    reduction (3-45, 90-135) prefers the larger value, then the larger original
 *    column index.  Begins are 0-based original indices (577-581, 589-601).
 */
#include "gbsw_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Residue order of the reference's 24x24 table (driver.cpp:252-258 maps each
 * of these letters to its position here; every other byte below 91 maps to 0). */
static const char AA_ORDER[] = "ARNDCQEGHILKMFPSTWYVBZX*";

int gbsw_oracle_encode_aa(int byte)
{
    if (byte < 0 || byte >= 91) return -1;
    for (int k = 0; k < 24; ++k)
        if (AA_ORDER[k] == byte) return k;
    return 0;
}

static uint32_t bitrev5(uint32_t x)
{
    uint32_t r = 0;
    for (int b = 0; b < 5; ++b) r |= ((x >> b) & 1u) << (4 - b);
    return r;
}

/* kernel.cpp:59-80: at each shfl_down step lane L keeps its own candidate on a
 * full tie, so within 32 lanes the survivor is the one whose lane index has the
 * smallest bit-reversal; kernel.cpp:157-179 repeats this over warp indices.
 * Columns >= 1024 do not exist in the reference (blockDim <= 1024); they are
 * ordered here by t>>10 first, then by the same two 5-bit levels. */
uint32_t gbsw_oracle_prio(uint32_t t)
{
    return ((t >> 10) << 10) | (bitrev5((t >> 5) & 31u) << 5) | bitrev5(t & 31u);
}

static int32_t imax(int32_t a, int32_t b) { return a > b ? a : b; }

typedef struct {
    int32_t* H;
    int32_t* F;
    int16_t* enc_r; /* AA: encoded row residues  */
    int16_t* enc_c; /* AA: encoded column residues */
    int32_t cap_c, cap_r;
} workspace;

static int ws_reserve(workspace* w, int32_t nr, int32_t nc)
{
    if (nc + 1 > w->cap_c) {
        free(w->H); free(w->F); free(w->enc_c);
        w->cap_c = nc + 1;
        w->H = (int32_t*)malloc(sizeof(int32_t) * (size_t)w->cap_c);
        w->F = (int32_t*)malloc(sizeof(int32_t) * (size_t)w->cap_c);
        w->enc_c = (int16_t*)malloc(sizeof(int16_t) * (size_t)w->cap_c);
        if (!w->H || !w->F || !w->enc_c) return -1;
    }
    if (nr + 1 > w->cap_r) {
        free(w->enc_r);
        w->cap_r = nr + 1;
        w->enc_r = (int16_t*)malloc(sizeof(int16_t) * (size_t)w->cap_r);
        if (!w->enc_r) return -1;
    }
    return 0;
}

static void ws_free(workspace* w)
{
    free(w->H); free(w->F); free(w->enc_r); free(w->enc_c);
    memset(w, 0, sizeof(*w));
}

/* Row byte of reversed-or-not row i (1-based) and column byte of column t
 * (0-based) are fetched through these accessors. */
typedef struct {
    const char* R;
    const char* C;
    int32_t nr, nc;
    int reversed; /* reverse pass: row i -> R[nr-i], column t -> C[nc-1-t] */
} grid;

static inline char row_byte(const grid* g, int32_t i) { return g->reversed ? g->R[g->nr - i] : g->R[i - 1]; }
static inline char col_byte(const grid* g, int32_t t) { return g->reversed ? g->C[g->nc - 1 - t] : g->C[t]; }

/* Fill the AA encodings; -1 on a residue the reference cannot index. */
static int encode_grid(const grid* g, const gbsw_oracle_scoring* sc, workspace* w)
{
    if (sc->kind != GBSW_ORACLE_AA) return 0;
    for (int32_t i = 1; i <= g->nr; ++i) {
        int e = gbsw_oracle_encode_aa((signed char)row_byte(g, i));
        if (e < 0) return -1;
        w->enc_r[i] = (int16_t)e;
    }
    for (int32_t t = 0; t < g->nc; ++t) {
        int e = gbsw_oracle_encode_aa((signed char)col_byte(g, t));
        if (e < 0) return -1;
        w->enc_c[t] = (int16_t)e;
    }
    return 0;
}

/* Runs the reference DP over the grid.  mode 0 tracks the forward winner
 * (max value, then min row, then min prio(column)); mode 1 tracks the reverse
 * winner (max value, then min column t, then that column's first row). */
static void run_dp(const grid* g, const gbsw_oracle_scoring* sc, workspace* w, int mode,
                   int32_t* best, int32_t* best_i, int32_t* best_t)
{
    const int32_t open = sc->gap_open, ext = sc->gap_extend;
    const int aa = sc->kind == GBSW_ORACLE_AA;
    int32_t* Hp = w->H;
    int32_t* Fp = w->F;
    for (int32_t j = 0; j <= g->nc; ++j) { Hp[j] = 0; Fp[j] = 0; }
    int32_t bv = 0, bi = 0, bt = 0;
    uint32_t bprio = 0;
    for (int32_t i = 1; i <= g->nr; ++i) {
        const char rb = row_byte(g, i);
        const int16_t* mrow = aa ? sc->matrix + 24 * w->enc_r[i] : NULL;
        int32_t hdiag = 0, hleft = 0, eleft = 0;
        for (int32_t j = 1; j <= g->nc; ++j) {
            const int32_t f = imax(Fp[j] + ext, Hp[j] + open);
            const int32_t e = (j == 1) ? 0 : imax(eleft + ext, hleft + open);
            int32_t s;
            if (aa) s = mrow[w->enc_c[j - 1]];
            else s = (rb == col_byte(g, j - 1)) ? sc->match : sc->mismatch;
            const int32_t d = hdiag + s;
            const int32_t h = imax(imax(d, f), imax(e, 0));
            hdiag = Hp[j];
            Hp[j] = h;
            Fp[j] = f;
            hleft = h;
            eleft = e;
            if (h > bv) {
                bv = h; bi = i; bt = j - 1;
                bprio = gbsw_oracle_prio((uint32_t)(j - 1));
            } else if (h == bv && h > 0) {
                if (mode == 0) {
                    /* rows arrive in order: only a same-row tie can win, on prio */
                    if (i == bi && gbsw_oracle_prio((uint32_t)(j - 1)) < bprio) {
                        bt = j - 1;
                        bprio = gbsw_oracle_prio((uint32_t)(j - 1));
                    }
                } else if (j - 1 < bt) {
                    /* a column's first hit comes first in row order */
                    bt = j - 1; bi = i;
                }
            }
        }
    }
    *best = bv; *best_i = bi; *best_t = bt;
}

static int forward_ws(const char* ref, int32_t lr, const char* q, int32_t lq,
                      const gbsw_oracle_scoring* sc, workspace* w, int32_t* score,
                      int32_t* ref_end, int32_t* query_end)
{
    grid g;
    const int swapped = lr < lq; /* kernel.cpp:248: rows are B when A is shorter */
    g.R = swapped ? q : ref; g.nr = swapped ? lq : lr;
    g.C = swapped ? ref : q; g.nc = swapped ? lr : lq;
    g.reversed = 0;
    if (ws_reserve(w, g.nr, g.nc)) return -2;
    if (encode_grid(&g, sc, w)) return -1;
    int32_t bv, bi, bt;
    run_dp(&g, sc, w, 0, &bv, &bi, &bt);
    if (bv <= 0) { *score = 0; *ref_end = 0; *query_end = 0; return 0; }
    *score = bv;
    if (swapped) { *query_end = bi; *ref_end = bt + 1; }   /* kernel.cpp:381-386 */
    else         { *ref_end = bi;   *query_end = bt + 1; } /* kernel.cpp:387-392 */
    return 0;
}

static int reverse_ws(const char* ref, int32_t ea, const char* q, int32_t eb,
                      const gbsw_oracle_scoring* sc, workspace* w, int32_t* rev_max,
                      int32_t* ref_begin, int32_t* query_begin)
{
    grid g;
    const int swapped = ea < eb; /* kernel.cpp:444 */
    g.R = swapped ? q : ref; g.nr = swapped ? eb : ea;
    g.C = swapped ? ref : q; g.nc = swapped ? ea : eb;
    g.reversed = 1;
    if (ws_reserve(w, g.nr, g.nc)) return -2;
    if (encode_grid(&g, sc, w)) return -1;
    int32_t bv, bi, bt;
    run_dp(&g, sc, w, 1, &bv, &bi, &bt);
    *rev_max = bv;
    if (bv <= 0) { *ref_begin = 0; *query_begin = 0; return 0; }
    const int32_t orow = g.nr - bi;      /* kernel.cpp:579 */
    const int32_t ocol = g.nc - bt - 1;  /* kernel.cpp:580 */
    if (swapped) { *query_begin = orow; *ref_begin = ocol; } /* 591-595 */
    else         { *ref_begin = orow; *query_begin = ocol; } /* 596-600 */
    return 0;
}

int gbsw_oracle_forward(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                        const gbsw_oracle_scoring* sc, int32_t* score, int32_t* ref_end,
                        int32_t* query_end)
{
    workspace w; memset(&w, 0, sizeof(w));
    int rc = forward_ws(ref, ref_len, query, query_len, sc, &w, score, ref_end, query_end);
    ws_free(&w);
    return rc;
}

int gbsw_oracle_reverse(const char* ref, int32_t ref_end, const char* query, int32_t query_end,
                        const gbsw_oracle_scoring* sc, int32_t* rev_max, int32_t* ref_begin,
                        int32_t* query_begin)
{
    workspace w; memset(&w, 0, sizeof(w));
    int rc = reverse_ws(ref, ref_end, query, query_end, sc, &w, rev_max, ref_begin, query_begin);
    ws_free(&w);
    return rc;
}

static int align_ws(const char* ref, int32_t lr, const char* q, int32_t lq,
                    const gbsw_oracle_scoring* sc, workspace* w, gbsw_oracle_result* out)
{
    memset(out, 0, sizeof(*out));
    int32_t s, ea, eb, rmax, ba, bb;
    int rc = forward_ws(ref, lr, q, lq, sc, w, &s, &ea, &eb);
    if (rc) return rc;
    if (s <= 0) return 0;
    rc = reverse_ws(ref, ea, q, eb, sc, w, &rmax, &ba, &bb);
    if (rc) return rc;
    out->score = s;
    out->ref_end = ea;
    out->query_end = eb;
    out->ref_begin = ba;
    out->query_begin = bb;
    return 0;
}

int gbsw_oracle_align(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                      const gbsw_oracle_scoring* sc, gbsw_oracle_result* out)
{
    workspace w; memset(&w, 0, sizeof(w));
    int rc = align_ws(ref, ref_len, query, query_len, sc, &w, out);
    ws_free(&w);
    return rc;
}

int gbsw_oracle_threads(int nthreads)
{
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    return nthreads;
#else
    (void)nthreads;
    return 1;
#endif
}

int64_t gbsw_oracle_align_batch(int64_t n, const char* ref, const uint64_t* ref_off,
                                const char* query, const uint64_t* query_off,
                                const gbsw_oracle_scoring* sc, gbsw_oracle_result* out,
                                int nthreads)
{
    int64_t first_bad = 0;
    const int nt = gbsw_oracle_threads(nthreads);
    (void)nt;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        workspace w; memset(&w, 0, sizeof(w));
        int64_t bad = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t k = 0; k < n; ++k) {
            const int rc = align_ws(ref + ref_off[k], (int32_t)(ref_off[k + 1] - ref_off[k]),
                                    query + query_off[k], (int32_t)(query_off[k + 1] - query_off[k]),
                                    sc, &w, &out[k]);
            if (rc && (bad == 0 || k + 1 < bad)) bad = k + 1;
        }
        ws_free(&w);
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (bad && (first_bad == 0 || bad < first_bad)) first_bad = bad;
        }
    }
    return first_bad ? -first_bad : 0;
}

/* ------------------------------------------------------------------ paths
 *
 * Alignment path of one pair: a capability the reference declares but never
 * implements (traceBack, kernel.hpp:36-39, no definition in kernel.cpp).  This
 * build defines it as the highest-scoring GLOBAL alignment of
 * ref[ref_begin, ref_end) against query[query_begin, query_end) (the rectangle
 * the forward and reverse passes report), scored like the forward pass: a run
 * of L gap positions scores gap_open + (L-1) gap_extend, and the substitution is
 * s(row byte, column byte) with rows = the longer full sequence (ref on ties,
 * kernel.cpp:248-259; the protein table is indexed enc(row)*24 + enc(col),
 * kernel.cpp:785-788).  Among optimal alignments the path is the one whose op
 * sequence, read from its end backwards, is lexicographically smallest with
 * M < D < I (M = a residue pair, D = a ref residue against a gap, I = a query
 * residue against a gap).
 *
 * This restatement keeps the full H/E/F matrices and decides each backward step
 * from their values (the device engine stores tightness bits instead). */

#define TB_NEG (-(1 << 29))

typedef struct {
    int32_t *H, *E, *F;
    uint8_t* ops;
    size_t cap_cells, cap_ops;
} tb_workspace;

static int tb_reserve(tb_workspace* w, size_t cells, size_t nops)
{
    if (cells > w->cap_cells) {
        free(w->H); free(w->E); free(w->F);
        w->cap_cells = cells;
        w->H = (int32_t*)malloc(sizeof(int32_t) * cells);
        w->E = (int32_t*)malloc(sizeof(int32_t) * cells);
        w->F = (int32_t*)malloc(sizeof(int32_t) * cells);
        if (!w->H || !w->E || !w->F) return -1;
    }
    if (nops > w->cap_ops) {
        free(w->ops);
        w->cap_ops = nops;
        w->ops = (uint8_t*)malloc(nops);
        if (!w->ops) return -1;
    }
    return 0;
}

static void tb_free(tb_workspace* w)
{
    free(w->H); free(w->E); free(w->F); free(w->ops);
    memset(w, 0, sizeof(*w));
}

static int32_t tb_sub(const gbsw_oracle_scoring* sc, char rowc, char colc)
{
    if (sc->kind == GBSW_ORACLE_AA)
        return sc->matrix[gbsw_oracle_encode_aa((signed char)rowc) * 24 + gbsw_oracle_encode_aa((signed char)colc)];
    return rowc == colc ? sc->match : sc->mismatch;
}

enum { TB_M = 0, TB_I = 1, TB_D = 2 }; /* BAM CIGAR op codes */

static int tb_path(const char* a, int32_t m, const char* b, int32_t n, int swap, const gbsw_oracle_scoring* sc,
                   tb_workspace* w, uint32_t* ops, int32_t cap, int32_t* nops, int32_t* path_score)
{
    const int32_t open = sc->gap_open, ext = sc->gap_extend;
    const size_t W = (size_t)n + 1;
    if (sc->kind == GBSW_ORACLE_AA) {
        for (int32_t i = 0; i < m; ++i) if (gbsw_oracle_encode_aa((signed char)a[i]) < 0) return -1;
        for (int32_t j = 0; j < n; ++j) if (gbsw_oracle_encode_aa((signed char)b[j]) < 0) return -1;
    }
    if (tb_reserve(w, ((size_t)m + 1) * W, (size_t)m + (size_t)n + 1)) return -2;
    int32_t *H = w->H, *E = w->E, *F = w->F;
#define S_(i, j) (swap ? tb_sub(sc, b[(j) - 1], a[(i) - 1]) : tb_sub(sc, a[(i) - 1], b[(j) - 1]))
    H[0] = 0; E[0] = F[0] = TB_NEG;
    for (int32_t j = 1; j <= n; ++j) { E[j] = open + (j - 1) * ext; H[j] = E[j]; F[j] = TB_NEG; }
    for (int32_t i = 1; i <= m; ++i) {
        const size_t r = (size_t)i * W, u = r - W;
        F[r] = open + (i - 1) * ext; H[r] = F[r]; E[r] = TB_NEG;
        for (int32_t j = 1; j <= n; ++j) {
            const int32_t e = imax(E[r + j - 1] + ext, H[r + j - 1] + open);
            const int32_t f = imax(F[u + j] + ext, H[u + j] + open);
            const int32_t d = H[u + j - 1] + S_(i, j);
            E[r + j] = e; F[r + j] = f;
            H[r + j] = imax(d, imax(e, f));
        }
    }
    *path_score = H[(size_t)m * W + n];
    /* backward walk: state 0 = H, 1 = F (in a D run), 2 = E (in an I run) */
    int32_t i = m, j = n, k = 0, st = 0;
#define AT(X, i, j) X[(size_t)(i) * W + (j)]
    while (i > 0 || j > 0) {
        if (i == 0) { w->ops[k++] = TB_I; --j; continue; }
        if (j == 0) { w->ops[k++] = TB_D; --i; continue; }
        if (st == 0) {
            if (AT(H, i, j) == AT(H, i - 1, j - 1) + S_(i, j)) { w->ops[k++] = TB_M; --i; --j; }
            else if (AT(H, i, j) == AT(F, i, j)) st = 1;
            else st = 2;
            continue;
        }
        if (st == 1) {
            w->ops[k++] = TB_D;
            const int open_t = AT(F, i, j) == AT(H, i - 1, j) + open;
            const int ext_t = AT(F, i, j) == AT(F, i - 1, j) + ext;
            const int m_next = open_t && i >= 2 && AT(H, i - 1, j) == AT(H, i - 2, j - 1) + S_(i - 1, j);
            st = m_next ? 0 : ext_t ? 1 : 0;
            --i;
            continue;
        }
        w->ops[k++] = TB_I;
        {
            const int open_t = AT(E, i, j) == AT(H, i, j - 1) + open;
            const int ext_t = AT(E, i, j) == AT(E, i, j - 1) + ext;
            const int md_next = open_t && j >= 2 &&
                                (AT(H, i, j - 1) == AT(H, i - 1, j - 2) + S_(i, j - 1) || AT(H, i, j - 1) == AT(F, i, j - 1));
            st = md_next ? 0 : ext_t ? 2 : 0;
            --j;
        }
    }
#undef AT
#undef S_
    /* forward order, run-length encoded as BAM ops (length << 4 | op) */
    int32_t c = 0;
    for (int32_t x = k - 1; x >= 0;) {
        int32_t y = x;
        while (y > 0 && w->ops[y - 1] == w->ops[x]) --y;
        if (c >= cap) return -3;
        ops[c++] = ((uint32_t)(x - y + 1) << 4) | w->ops[x];
        x = y - 1;
    }
    *nops = c;
    return 0;
}

int gbsw_oracle_traceback(const char* ref, int32_t ref_len, const char* query, int32_t query_len,
                          const gbsw_oracle_scoring* sc, const gbsw_oracle_result* res, uint32_t* ops,
                          int32_t cap, int32_t* nops, int32_t* path_score)
{
    tb_workspace w; memset(&w, 0, sizeof(w));
    *nops = 0; *path_score = 0;
    int rc = 0;
    if (res->score > 0)
        rc = tb_path(ref + res->ref_begin, res->ref_end - res->ref_begin, query + res->query_begin,
                     res->query_end - res->query_begin, ref_len < query_len, sc, &w, ops, cap, nops, path_score);
    tb_free(&w);
    return rc;
}

int64_t gbsw_oracle_traceback_batch(int64_t n, const char* ref, const uint64_t* ref_off, const char* query,
                                    const uint64_t* query_off, const gbsw_oracle_scoring* sc,
                                    const gbsw_oracle_result* res, uint32_t* ops, const int64_t* cap_off,
                                    int32_t* nops, int32_t* path_score, int nthreads)
{
    int64_t first_bad = 0;
    const int nt = gbsw_oracle_threads(nthreads);
    (void)nt;
#ifdef _OPENMP
#pragma omp parallel num_threads(nt)
#endif
    {
        tb_workspace w; memset(&w, 0, sizeof(w));
        int64_t bad = 0;
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
        for (int64_t k = 0; k < n; ++k) {
            const gbsw_oracle_result* r = &res[k];
            const int32_t lr = (int32_t)(ref_off[k + 1] - ref_off[k]), lq = (int32_t)(query_off[k + 1] - query_off[k]);
            nops[k] = 0; path_score[k] = 0;
            int rc = 0;
            if (r->score > 0)
                rc = tb_path(ref + ref_off[k] + r->ref_begin, r->ref_end - r->ref_begin,
                             query + query_off[k] + r->query_begin, r->query_end - r->query_begin, lr < lq, sc, &w,
                             ops + cap_off[k], (int32_t)(cap_off[k + 1] - cap_off[k]), &nops[k], &path_score[k]);
            if (rc && (bad == 0 || k + 1 < bad)) bad = k + 1;
        }
        tb_free(&w);
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            if (bad && (first_bad == 0 || bad < first_bad)) first_bad = bad;
        }
    }
    return first_bad ? -first_bad : 0;
}
