/* Host sanitizer self-test of the oracle: runs the hand fixtures and a seeded
 * random batch under ASan/UBSan (make -C oracle asan).  Test infrastructure. */
#include "gbsw_oracle.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int check(const char* name, int got, int want)
{
    if (got != want) { printf("FAIL %s: got %d want %d\n", name, got, want); return 1; }
    return 0;
}

int main(void)
{
    int bad = 0;
    gbsw_oracle_scoring sc;
    memset(&sc, 0, sizeof(sc));
    sc.kind = GBSW_ORACLE_DNA; sc.match = 1; sc.mismatch = -3; sc.gap_open = -3; sc.gap_extend = -1;
    gbsw_oracle_result r;
    /* Full tie in row 9: column t=8 beats t=3 on bit-reversed priority. */
    const char* ref = "GGGGGACGTGGGG";
    const char* qry = "ACGTTACGT";
    gbsw_oracle_align(ref, 13, qry, 9, &sc, &r);
    bad |= check("score", r.score, 4);
    bad |= check("ref_end", r.ref_end, 9);
    bad |= check("query_end", r.query_end, 9);
    bad |= check("ref_begin", r.ref_begin, 5);
    bad |= check("query_begin", r.query_begin, 5);

    /* seeded random batch through the OpenMP driver */
    enum { N = 200 };
    uint64_t ro[N + 1], qo[N + 1];
    char* rs = (char*)malloc(N * 64);
    char* qs = (char*)malloc(N * 64);
    unsigned s = 12345u;
    ro[0] = qo[0] = 0;
    for (int k = 0; k < N; ++k) {
        int lr = (int)((s = s * 1103515245u + 12345u) >> 16) % 60;
        int lq = (int)((s = s * 1103515245u + 12345u) >> 16) % 60;
        for (int i = 0; i < lr; ++i) rs[ro[k] + i] = "ACGT"[((s = s * 1103515245u + 12345u) >> 16) & 3];
        for (int i = 0; i < lq; ++i) qs[qo[k] + i] = "ACGT"[((s = s * 1103515245u + 12345u) >> 16) & 3];
        ro[k + 1] = ro[k] + lr;
        qo[k + 1] = qo[k] + lq;
    }
    gbsw_oracle_result* out = (gbsw_oracle_result*)calloc(N, sizeof(gbsw_oracle_result));
    bad |= check("batch rc", (int)gbsw_oracle_align_batch(N, rs, ro, qs, qo, &sc, out, 2), 0);
    free(out); free(rs); free(qs);
    printf(bad ? "oracle selftest FAILED\n" : "oracle selftest ok\n");
    return bad;
}
