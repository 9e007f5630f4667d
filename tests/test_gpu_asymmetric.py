"""GPU parity with asymmetric substitution tables (the reference accepts any
576 entries, driver.cpp:304-308): the forward pass scores
M[enc(row)*24 + enc(col)] with rows the longer sequence (kernel.cpp:785-788);
the reverse pass takes its rows from the end coordinates (kernel.cpp:444-459,
1021-1030), so when that swaps rows and columns its maximum is not the forward
score and the begins follow its own maximum (kernel.cpp:3-45, 90-135).  Every
output equals the oracle's, which the literal kernel model pins on such tables
(tests/test_oracle.py)."""
import random

import numpy as np
import pytest

import oracle_lib as ol
import gbsw
from gbsw import synth
from gbsw.scoring import BLOSUM62
from test_oracle import _asym_table

pytestmark = pytest.mark.gpu

FIELDS = gbsw.FIELDS


@pytest.fixture(scope="module")
def ctx():
    c = gbsw.Context(0)
    yield c
    c.close()


def assert_same(got, exp, label):
    for f in FIELDS:
        bad = np.nonzero(got[f] != exp[f])[0]
        if len(bad):
            k = int(bad[0])
            detail = {g: (int(got[g][k]), int(exp[g][k])) for g in FIELDS}
            raise AssertionError("%s: %d pairs differ in %s; first pair %d (got, want) %s" %
                                 (label, len(bad), f, k, detail))


def skewed_blosum(seed):
    """BLOSUM62 with its upper triangle shifted: M[a][b] != M[b][a] above the diagonal."""
    rng = random.Random(seed)
    m = list(BLOSUM62)
    for r in range(24):
        for c in range(r + 1, 24):
            m[r * 24 + c] += rng.choice([-2, -1, 1, 2])
    return m


def check(ctx, rb, ro, qb, qo, table, go, ge, label, modes=("plain", "batched")):
    exp = ol.align_batch(rb, ro, qb, qo, ol.aa_scoring(table, go, ge))
    for mode in modes:
        print("case", label, mode, len(ro) - 1, flush=True)
        got = ctx.align_packed(rb, ro, qb, qo, gbsw.protein_scoring(table, go, ge), mode)
        assert_same(got, exp, "%s/%s" % (label, mode))
    return exp


def one_table_cases(seed, n):
    """Pairs for one random asymmetric table: most end in a mutated copy of a
    stretch at the ref's start with filler residues (negative against the
    rest) around it, so the reverse pass swaps rows and columns."""
    rng = random.Random(seed)
    filler = "WYF"
    table = _asym_table(rng, filler)
    core_alpha = "ARNDCQEGHILKMPSTV"
    refs, qrys = [], []
    for _ in range(n):
        alpha = "".join(rng.sample(core_alpha, rng.randint(2, 5)))
        if rng.random() < 0.7:
            core = "".join(rng.choice(alpha) for _ in range(rng.randint(6, 40)))
            copy = "".join(rng.choice(alpha) if rng.random() < 0.25 else x for x in core)
            refs.append(core + "".join(rng.choice(filler) for _ in range(rng.randint(10, 300))))
            qrys.append("".join(rng.choice(filler) for _ in range(rng.randint(4, 60))) + copy)
        else:
            refs.append("".join(rng.choice(alpha) for _ in range(rng.randint(1, 300))))
            qrys.append("".join(rng.choice(alpha) for _ in range(rng.randint(1, 300))))
    return table, refs, qrys


@pytest.mark.parametrize("seed", [51, 52, 53])
def test_random_tables_swapped_reverse_orientation(ctx, seed):
    """1,500 pairs per random asymmetric table, most with the reverse
    orientation swapped and its maximum unlike the forward score."""
    table, refs, qrys = one_table_cases(seed, 1500)
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = check(ctx, rb, ro, qb, qo, table, -4, -1, "asym random table %d" % seed)
    la, lq = np.diff(ro).astype(np.int64), np.diff(qo).astype(np.int64)
    swapped = (la >= lq) != (exp["ref_end"] >= exp["query_end"])
    assert int(np.sum(swapped & (exp["score"] > 0))) > 300


def test_protein_workload_with_a_skewed_blosum(ctx):
    """The 20k protein generator (4,000 pairs) under BLOSUM62 with a skewed upper
    triangle, gap -6/-1; the f16 option on and off."""
    w = synth.plain_protein(4000, seed=0x5EED0701)
    table = skewed_blosum(7)
    exp = check(ctx, w.ref, w.ref_off, w.query, w.query_off, table, -6, -1, "skewed blosum 4k")
    try:
        ctx.set_option(gbsw.OPT_F16, 0)
        got = ctx.align_packed(w.ref, w.ref_off, w.query, w.query_off, gbsw.protein_scoring(table, -6, -1), "batched")
        assert_same(got, exp, "skewed blosum f16 off")
    finally:
        ctx.set_option(gbsw.OPT_F16, 1)


def test_long_proteins_with_a_skewed_blosum(ctx):
    """Multi-strip pairs (refs <= 3000, queries <= 1500; the split form is not
    used with asymmetric tables) and long identical W-rich cores whose scores
    pass the int16 range (recomputed with int32 cells)."""
    w = synth.long_protein(24, seed=0x5EED0702)
    refs = [bytes(w.pair(k)[0][:3000]) for k in range(w.n)]
    qrys = [bytes(w.pair(k)[1][:1500]) for k in range(w.n)]
    rng = random.Random(19)
    for k in range(4):
        core = "".join(rng.choice("WWWWC") for _ in range(rng.randint(3000, 3600))).encode()
        refs.append(core + b"AKR" * 40)
        qrys.append(b"GG" * 30 + core[: len(core) - rng.randint(0, 40)])
    rb, ro = gbsw.pack(refs)
    qb, qo = gbsw.pack(qrys)
    exp = check(ctx, rb, ro, qb, qo, skewed_blosum(8), -6, -1, "skewed blosum long")
    assert exp["score"].max() > 32767


def test_asymmetric_table_alignment_paths(ctx):
    """Alignment paths (gbsw_batch_traceback) of pairs aligned under an
    asymmetric table equal the oracle's (rows the longer full sequence)."""
    w = synth.plain_protein(600, seed=0x5EED0703)
    table = skewed_blosum(9)
    gsc, osc = gbsw.protein_scoring(table, -6, -1), ol.aa_scoring(table, -6, -1)
    exp = ol.align_batch(w.ref, w.ref_off, w.query, w.query_off, osc)
    b = ctx.batch(gbsw.PROTEIN, w.ref, w.ref_off, w.query, w.query_off)
    try:
        b.align(gsc, "batched")
        res = b.results()
        assert_same(res, exp, "asym paths: alignment")
        words, off, ps, _ = b.traceback(gsc)
        ew, eoff, eps = ol.traceback_batch(w.ref, w.ref_off, w.query, w.query_off, osc, exp)
        assert np.array_equal(off, eoff) and np.array_equal(words, ew) and np.array_equal(ps, eps)
    finally:
        b.close()
