"""GPU alignment paths (gbsw_batch_traceback) against the oracle's.

The path of a pair is defined in oracle/gbsw_oracle.c (best global alignment of
the rectangle the alignment reports, ties by the op sequence read from its end,
M < D < I) and pinned there by brute force (tests/test_traceback_oracle.py).
Bar: the same CIGAR words and path score for every pair.
"""
import random

import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = gbsw.Context(0)
    yield c
    c.close()


def pack(pairs):
    refs = [r.encode() if isinstance(r, str) else r for r, _ in pairs]
    qs = [q.encode() if isinstance(q, str) else q for _, q in pairs]
    ro = np.zeros(len(pairs) + 1, np.uint64)
    qo = np.zeros(len(pairs) + 1, np.uint64)
    np.cumsum([len(r) for r in refs], out=ro[1:])
    np.cumsum([len(q) for q in qs], out=qo[1:])
    rb = np.frombuffer(b"".join(refs), dtype=np.uint8).copy()
    qb = np.frombuffer(b"".join(qs), dtype=np.uint8).copy()
    return rb, ro, qb, qo


def check_paths(ctx, alphabet, rb, ro, qb, qo, gsc, osc, label, mode="batched"):
    print("case", label, mode, "pairs", len(ro) - 1, flush=True)
    b = ctx.batch(alphabet, rb, ro, qb, qo)
    try:
        b.align(gsc, mode)
        got = b.results()
        words, off, ps, ms = b.traceback(gsc)
    finally:
        b.close()
    res = np.zeros(len(ro) - 1, dtype=ol.RESULT_DTYPE)
    for f in gbsw.FIELDS:
        res[f] = got[f]
    exp = ol.align_batch(rb, ro, qb, qo, osc)
    for f in gbsw.FIELDS:
        assert np.array_equal(exp[f], got[f]), (label, f)
    ew, eoff, eps = ol.traceback_batch(rb, ro, qb, qo, osc, res)
    bad = np.nonzero((np.diff(off) != np.diff(eoff)) | (ps != eps))[0]
    if len(bad) == 0 and not np.array_equal(words, ew):
        k = int(np.searchsorted(eoff, int(np.nonzero(words != ew)[0][0]), side="right") - 1)
        bad = np.array([k])
    if len(bad):
        k = int(bad[0])
        raise AssertionError("%s: %d pairs differ; pair %d got %s (%d) want %s (%d)" % (
            label, len(bad), k, gbsw.cigar_string(words[off[k]:off[k + 1]]), ps[k],
            ol.cigar_string(ew[eoff[k]:eoff[k + 1]]), eps[k]))
    assert ms >= 0
    return words, off, ps, res


def test_dna_paths_equal_the_oracle(ctx):
    w = synth.plain_dna(3000, seed=21)
    check_paths(ctx, gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off, gbsw.dna_scoring(*w.scoring),
                ol.dna_scoring(*w.scoring), "plain_dna 3000")


def test_protein_paths_equal_the_oracle(ctx):
    w = synth.plain_protein(1500, seed=22)
    check_paths(ctx, gbsw.PROTEIN, w.ref, w.ref_off, w.query, w.query_off, gbsw.protein_scoring(BLOSUM62, *w.scoring),
                ol.aa_scoring(BLOSUM62, *w.scoring), "plain_protein 1500")


def test_long_and_uneven_dna_paths(ctx):
    # the DNA mix's length range: rectangles of up to ~1000 columns (16 stripes)
    w = synth.batched_dna(4000, seed=23)
    words, off, ps, res = check_paths(ctx, gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off,
                                      gbsw.dna_scoring(*w.scoring), ol.dna_scoring(*w.scoring), "batched_dna 4000")
    cols = res["query_end"] - res["query_begin"]
    assert cols.max() > 640


def test_long_rectangles(ctx):
    # rectangles of up to ~4,000 x 4,000 cells: tens of stripes, many ring refills
    rng = np.random.default_rng(31)
    pairs = []
    for k in range(24):
        n = int(rng.integers(1500, 4000))
        ref = rng.choice(np.frombuffer(b"ACGT", np.uint8), n)
        q = ref[int(rng.integers(0, 200)):].copy()
        mut = rng.random(len(q)) < 0.08
        q[mut] = rng.choice(np.frombuffer(b"ACGT", np.uint8), int(mut.sum()))
        if k % 3 == 0:  # a few indels
            cut = int(rng.integers(10, len(q) - 10))
            q = np.concatenate([q[:cut], q[cut + 3:]])
        pairs.append((ref.tobytes(), q.tobytes()))
    rb, ro, qb, qo = pack(pairs)
    words, off, ps, res = check_paths(ctx, gbsw.DNA, rb, ro, qb, qo, gbsw.dna_scoring(3, -3, -6, -1),
                                      ol.dna_scoring(3, -3, -6, -1), "long rectangles")
    assert (res["query_end"] - res["query_begin"]).max() > 1400


def test_long_protein_paths(ctx):
    # the long-protein generator's lengths (queries 1,024-4,096, refs 4,096-16,384)
    # on a few pairs: int8-profile alignments, rectangles of thousands of columns
    w = synth.long_protein(12, seed=0x5EED0401)
    check_paths(ctx, gbsw.PROTEIN, w.ref, w.ref_off, w.query, w.query_off, gbsw.protein_scoring(BLOSUM62, *w.scoring),
                ol.aa_scoring(BLOSUM62, *w.scoring), "long_protein 12")


def test_edge_cases(ctx):
    rng = random.Random(9)
    pairs = [
        ("A", "A"), ("A", "C"), ("ACGT", "TTTT"), ("GGGG", "CCCC"),  # tiny, score 0
        ("ACGTACGTAAAC", "ACGTTTACGTAAC"), ("GGGGGACGTGGGG", "ACGTTACGT"),
        ("ACGT" * 40, "ACGT" * 40),  # identical, 160 columns
        ("A" * 200, "A" * 70),  # repeats: many ties
        ("AC" * 150, "CA" * 120),
    ]
    for _ in range(200):  # ref shorter than query as often as not
        a = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 140)))
        b = "".join(rng.choice("ACGT") for _ in range(rng.randint(1, 140)))
        pairs.append((a, b))
    rb, ro, qb, qo = pack(pairs)
    for sc in [(3, -3, -6, -1), (1, -1, -1, -1), (2, -1, -3, -3)]:
        check_paths(ctx, gbsw.DNA, rb, ro, qb, qo, gbsw.dna_scoring(*sc), ol.dna_scoring(*sc), "edge %s" % (sc,))


def test_protein_paths_with_another_symmetric_table(ctx):
    rng = random.Random(4)
    mat = [0] * 576
    for x in range(24):
        for y in range(x + 1):
            mat[x * 24 + y] = mat[y * 24 + x] = 6 if x == y else rng.randint(-4, 2)
    letters = "ARNDCQEGHILKMFPSTWYV"
    pairs = []
    for _ in range(300):
        a = "".join(rng.choice(letters[:6]) for _ in range(rng.randint(5, 150)))
        b = "".join(rng.choice(letters[:6]) for _ in range(rng.randint(5, 150)))
        pairs.append((a, b))
    rb, ro, qb, qo = pack(pairs)
    check_paths(ctx, gbsw.PROTEIN, rb, ro, qb, qo, gbsw.protein_scoring(mat, -4, -1), ol.aa_scoring(mat, -4, -1),
                "protein symmetric table")


def test_plain_form_results_give_the_same_paths(ctx):
    w = synth.plain_dna(800, seed=24)
    check_paths(ctx, gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off, gbsw.dna_scoring(*w.scoring),
                ol.dna_scoring(*w.scoring), "plain_dna 800", mode="plain")


def test_small_capacity_reports_the_layout(ctx):
    w = synth.plain_dna(200, seed=25)
    sc = gbsw.dna_scoring(*w.scoring)
    b = ctx.batch(gbsw.DNA, w.ref, w.ref_off, w.query, w.query_off)
    try:
        b.align(sc, "batched")
        words, off, ps, _ = b.traceback(sc)
        with pytest.raises(gbsw.GbswError) as e:
            b.traceback(sc, capacity=int(off[-1]) - 1)
        assert e.value.code == gbsw.E_RANGE
        words2, off2, ps2, _ = b.traceback(sc, capacity=int(off[-1]))
        assert np.array_equal(words, words2) and np.array_equal(off, off2) and np.array_equal(ps, ps2)
    finally:
        b.close()
